// Analytic Jacobians of straight-line programs by forward-mode differentiation of the linear form.
//
// Each value that depends on the chosen inputs gets a tangent: a short list of (key, register) pairs, a key being one
// of the inputs differentiated against or a "hub". Tangents follow the usual rules (a product's tangent is
// b·da + a·db, an exponential's e^a·da, a select selects its operands' tangents, min/max select the smaller or larger
// operand's ...). A value whose tangent grows past kHubSize entries becomes a hub: its tangent is recorded once and the
// value then carries the single entry (hub, 1), so that everything computed from it (a third-body concentration,
// the bulk concentration, a running sum of rates) keeps a short tangent. At the differentiated outputs the hubs are
// expanded by the chain rule, latest first: d out / d x = Σ_h (d out / d h)(d h / d x) + the direct terms.
// Instructions are emitted in program order, each value's tangent right after the value, so a one-lane-per-state
// kernel keeps the original program's locality.
#include "jacobian.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <unordered_map>

namespace combustion::ast {

namespace {

constexpr size_t kHubSize = 6;

using Tangent = std::vector<std::pair<int32_t, int32_t>>;  // (key, register), sorted by key

class Builder {
 public:
  explicit Builder(int32_t n_input) {
    out_.n_input = n_input;
    next_ = n_input;
  }
  int32_t emit(LOp op, int32_t a = -1, int32_t b = -1, int32_t c = -1, double k = 0) {
    out_.code.push_back({op, next_, a, b, c, k});
    return next_++;
  }
  int32_t constant(double k) {
    uint64_t bits;
    std::memcpy(&bits, &k, 8);
    auto it = constants_.find(bits);
    if (it != constants_.end()) return it->second;
    const int32_t r = emit(LOp::Const, -1, -1, -1, k);
    constants_.emplace(bits, r);
    value_[r] = k;
    return r;
  }
  bool is_constant(int32_t r, double k) const {
    auto it = value_.find(r);
    return it != value_.end() && it->second == k;
  }
  int32_t neg(int32_t a) {
    auto it = value_.find(a);
    if (it != value_.end()) return constant(-it->second);
    return emit(LOp::Neg, a);
  }
  int32_t add(int32_t a, int32_t b) {
    if (is_constant(a, 0.)) return b;
    if (is_constant(b, 0.)) return a;
    return emit(LOp::Add, a, b);
  }
  int32_t sub(int32_t a, int32_t b) {
    if (is_constant(b, 0.)) return a;
    if (is_constant(a, 0.)) return neg(b);
    return emit(LOp::Sub, a, b);
  }
  int32_t mul(int32_t a, int32_t b) {
    if (is_constant(a, 1.)) return b;
    if (is_constant(b, 1.)) return a;
    if (is_constant(a, -1.)) return neg(b);
    if (is_constant(b, -1.)) return neg(a);
    if (is_constant(a, 0.) || is_constant(b, 0.)) return constant(0.);
    return emit(LOp::Mul, a, b);
  }
  Linear& linear() { return out_; }
  int32_t next() const { return next_; }

 private:
  Linear out_;
  int32_t next_;
  std::unordered_map<uint64_t, int32_t> constants_;
  std::unordered_map<int32_t, double> value_;
};

// a·x + b·y over the union of keys (absent entries are zero); coefficients are registers, -1 for 1.
Tangent combine(Builder& B, const Tangent& x, int32_t a, const Tangent& y, int32_t b) {
  Tangent out;
  size_t i = 0, j = 0;
  auto scale = [&](int32_t c, int32_t t) { return c < 0 ? t : B.mul(c, t); };
  while (i < x.size() || j < y.size()) {
    if (j == y.size() || (i < x.size() && x[i].first < y[j].first)) {
      out.push_back({x[i].first, scale(a, x[i].second)});
      i++;
    } else if (i == x.size() || y[j].first < x[i].first) {
      out.push_back({y[j].first, scale(b, y[j].second)});
      j++;
    } else {
      out.push_back({x[i].first, B.add(scale(a, x[i].second), scale(b, y[j].second))});
      i++;
      j++;
    }
  }
  return out;
}

Tangent scaled(Builder& B, const Tangent& x, int32_t c) {
  Tangent out;
  for (const auto& e : x) out.push_back({e.first, B.mul(c, e.second)});
  return out;
}

// select(cond, x, y) entrywise over the union of keys.
Tangent selected(Builder& B, int32_t cond, const Tangent& x, const Tangent& y) {
  Tangent out;
  size_t i = 0, j = 0;
  const int32_t zero = B.constant(0.);
  while (i < x.size() || j < y.size()) {
    int32_t key, tx = zero, ty = zero;
    if (j == y.size() || (i < x.size() && x[i].first < y[j].first)) {
      key = x[i].first;
      tx = x[i++].second;
    } else if (i == x.size() || y[j].first < x[i].first) {
      key = y[j].first;
      ty = y[j++].second;
    } else {
      key = x[i].first;
      tx = x[i++].second;
      ty = y[j++].second;
    }
    out.push_back({key, tx == ty ? tx : B.emit(LOp::Select, cond, tx, ty)});
  }
  return out;
}

}  // namespace

Linear jacobian(const Linear& l, const std::vector<int32_t>& wrt, const std::vector<int32_t>& rows) {
  const int32_t n_wrt = (int32_t)wrt.size();
  std::vector<int32_t> key_of_input(l.n_input, -1);
  for (int32_t d = 0; d < n_wrt; d++) {
    if (wrt[d] < 0 || wrt[d] >= l.n_input) throw Error("jacobian: an input to differentiate against is out of range");
    if (key_of_input[wrt[d]] >= 0) throw Error("jacobian: an input is listed twice");
    key_of_input[wrt[d]] = d;
  }
  for (int32_t o : rows)
    if (o < 0 || o >= (int32_t)l.output.size()) throw Error("jacobian: an output to differentiate is out of range");

  // Which registers depend on a differentiated input, and which reach a differentiated output.
  std::vector<char> depends(l.n_registers, 0), needed(l.n_registers, 0);
  for (int32_t i = 0; i < l.n_input; i++) depends[i] = key_of_input[i] >= 0;
  auto operands = [](const Instruction& i, int32_t* ops) {
    int n = 0;
    switch (i.op) {
      case LOp::Const: break;
      case LOp::Neg: case LOp::Sqrt: case LOp::Exp: case LOp::Ln: case LOp::Sq: ops[n++] = i.a; break;
      case LOp::Select: ops[n++] = i.b; ops[n++] = i.c; break;  // the condition carries no derivative
      default: ops[n++] = i.a; ops[n++] = i.b; break;
    }
    return n;
  };
  for (const Instruction& i : l.code) {
    int32_t ops[3];
    const int n = operands(i, ops);
    if (i.op == LOp::LessOrEqual) continue;
    for (int k = 0; k < n; k++) depends[i.dst] = depends[i.dst] || depends[ops[k]];
  }
  for (int32_t o : rows) needed[l.output[o]] = 1;
  for (auto it = l.code.rbegin(); it != l.code.rend(); ++it) {
    if (!needed[it->dst] || it->op == LOp::LessOrEqual) continue;
    int32_t ops[3];
    const int n = operands(*it, ops);
    for (int k = 0; k < n; k++) needed[ops[k]] = 1;
  }

  Builder B(l.n_input);
  std::vector<int32_t> reg(l.n_registers, -1);
  for (int32_t i = 0; i < l.n_input; i++) reg[i] = i;
  std::vector<Tangent> tangent(l.n_registers);
  std::vector<Tangent> hubs;  // hub h's tangent, keys < n_wrt + h
  const int32_t one = B.constant(1.);
  for (int32_t i = 0; i < l.n_input; i++)
    if (key_of_input[i] >= 0) tangent[i] = {{key_of_input[i], one}};

  for (const Instruction& i : l.code) {
    auto R = [&](int32_t r) { return reg[r]; };
    int32_t d;
    if (i.op == LOp::Const)
      d = B.constant(i.k);
    else
      d = B.emit(i.op, i.a >= 0 ? R(i.a) : -1, i.b >= 0 && i.op != LOp::Neg && i.op != LOp::Sqrt && i.op != LOp::Exp && i.op != LOp::Ln &&
                                                   i.op != LOp::Sq ? R(i.b) : -1,
                 i.op == LOp::Select ? R(i.c) : -1, i.k);
    reg[i.dst] = d;
    if (!depends[i.dst] || !needed[i.dst]) continue;
    const Tangent empty;
    auto T = [&](int32_t r) -> const Tangent& { return depends[r] ? tangent[r] : empty; };
    Tangent t;
    switch (i.op) {
      case LOp::Neg: t = scaled(B, T(i.a), B.constant(-1.)); break;
      case LOp::Add: t = combine(B, T(i.a), -1, T(i.b), -1); break;
      case LOp::Sub: t = combine(B, T(i.a), -1, T(i.b), B.constant(-1.)); break;
      case LOp::Mul: t = combine(B, T(i.a), R(i.b), T(i.b), R(i.a)); break;
      case LOp::Div: {  // (da - q db) / b with q = a / b
        const int32_t inv = B.emit(LOp::Div, one, R(i.b));
        const int32_t mq = B.neg(B.mul(d, inv));
        t = combine(B, T(i.a), inv, T(i.b), mq);
        break;
      }
      case LOp::Sq: t = scaled(B, T(i.a), B.mul(B.constant(2.), R(i.a))); break;
      case LOp::Sqrt: t = scaled(B, T(i.a), B.emit(LOp::Div, B.constant(0.5), d)); break;
      case LOp::Exp: t = scaled(B, T(i.a), d); break;
      case LOp::Ln: t = scaled(B, T(i.a), B.emit(LOp::Div, one, R(i.a))); break;
      case LOp::Min: t = selected(B, B.emit(LOp::LessOrEqual, R(i.a), R(i.b)), T(i.a), T(i.b)); break;
      case LOp::Max: t = selected(B, B.emit(LOp::LessOrEqual, R(i.b), R(i.a)), T(i.a), T(i.b)); break;
      case LOp::Select: t = selected(B, R(i.a), T(i.b), T(i.c)); break;
      default: throw Error("jacobian: unexpected instruction");
    }
    // drop entries that folded to the constant 0
    t.erase(std::remove_if(t.begin(), t.end(), [&](const std::pair<int32_t, int32_t>& e) { return B.is_constant(e.second, 0.); }), t.end());
    if (t.size() > kHubSize) {
      hubs.push_back(std::move(t));
      t = {{n_wrt + (int32_t)hubs.size() - 1, one}};
    }
    tangent[i.dst] = std::move(t);
  }

  // Outputs: the program's own, then d row / d wrt[j] for each differentiated row, with the hubs expanded.
  std::vector<int32_t> outputs;
  for (int32_t r : l.output) outputs.push_back(reg[r]);
  const int32_t zero = B.constant(0.);
  for (int32_t o : rows) {
    const int32_t r = l.output[o];
    std::map<int32_t, int32_t> acc;  // key -> register
    if (depends[r])
      for (const auto& e : tangent[r]) acc[e.first] = e.second;
    while (!acc.empty() && acc.rbegin()->first >= n_wrt) {
      const auto last = std::prev(acc.end());
      const int32_t h = last->first - n_wrt, c = last->second;
      acc.erase(last);
      for (const auto& e : hubs[h]) {
        const int32_t term = B.mul(c, e.second);
        auto it = acc.find(e.first);
        if (it == acc.end())
          acc.emplace(e.first, term);
        else
          it->second = B.add(it->second, term);
      }
    }
    for (int32_t dkey = 0; dkey < n_wrt; dkey++) {
      auto it = acc.find(dkey);
      outputs.push_back(it == acc.end() ? zero : it->second);
    }
  }
  Linear& out = B.linear();
  out.n_registers = B.next();
  out.output = std::move(outputs);
  return out;
}

}  // namespace combustion::ast
