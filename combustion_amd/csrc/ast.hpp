// Per-state straight-line program: the IR the builders produce and the emitters consume.
// Mirrors ast/lib.rs: Expr operators (:30-60), Statement::{Value, Select} (:139-143), Function (:328-333),
// the constant-folding rules of add/sub/mul/div (:195-232), Σ (:337-343), exp/ln (:373-382), and the
// common-subexpression reuse of reaction::Block::def (reaction.rs:9-22). Expression nodes are hash-consed,
// so structurally equal expressions share one node index and CSE is a lookup on that index.
#pragma once
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace combustion::ast {

enum class Op : uint8_t { Const, Value, Neg, Min, Max, Add, Sub, LessOrEqual, Mul, Div, Sqrt, Exp, Ln, Sq };
const char* op_name(Op op);

struct Node {
  Op op;
  int32_t a = -1, b = -1;  // operands (node indices); for Value, a is the value id
  double c = 0;            // Const payload
};
using E = int32_t;  // expression = node index

struct Statement {
  bool select = false;
  int32_t id = -1;  // Def: defined value
  E value = -1;     // Def: expression
  E condition = -1; // Select
  std::vector<E> t, f;
  std::vector<int32_t> results;
};

struct Function {
  int32_t n_input = 0;
  std::vector<Node> nodes;
  std::vector<Statement> statements;
  std::vector<E> output;
  std::vector<std::string> values;  // one name per value id; inputs first
};

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Block {
 public:
  explicit Block(std::vector<std::string> input_names);
  Function& function() { return f_; }
  Function finish(std::vector<E> outputs);

  E c(double x);           // constant; like ast::f64, it must be finite and within f32 range
  E v(int32_t id);         // value reference
  std::optional<double> f64(E e) const;
  bool is_leaf(E e) const;
  const Node& node(E e) const { return f_.nodes[e]; }

  // ast/lib.rs:179-246 (folding rules included)
  E neg(E x);
  E add(E a, E b);
  E sub(E a, E b);
  E mul(E a, E b);
  E div(E a, E b);
  E min(E a, E b);
  E max(E a, E b);
  E le(E a, E b);
  E sqrt(E x);
  E sq(E x);
  E exp(E x);
  E ln(E x);                    // ln(x0, x): constant operands fold to std::log
  std::optional<E> sigma(const std::vector<E>& terms);  // Σ: drops zero constants, left fold
  E sum(const std::vector<E>& terms);

  int32_t value(const std::string& name);  // new value id (Select result)
  int32_t def(E e, const std::string& name);  // reaction::Block::def: reuses an identical earlier definition
  E defv(E e, const std::string& name) { return v(def(e, name)); }
  E edef(E e, const std::string& name) { return is_leaf(e) ? e : defv(e, name); }
  void select(E condition, std::vector<E> t, std::vector<E> f, std::vector<int32_t> results);

 private:
  Function f_;
  std::unordered_map<uint64_t, std::vector<E>> intern_;
  std::unordered_map<E, int32_t> defined_;
  E make(Op op, E a, E b, double c = 0);
  bool is_f32(E e, float x) const;
};

// Linear SSA form shared by the HIP emitter and the program export: every node that a statement or output
// evaluates becomes one instruction on a register; Select becomes one Select instruction per result.
enum class LOp : uint8_t { Const, Neg, Min, Max, Add, Sub, LessOrEqual, Mul, Div, Sqrt, Exp, Ln, Sq, Select };
struct Instruction {
  LOp op;
  int32_t dst, a, b, c;  // registers (c: Select's false operand)
  double k;              // Const payload
};
struct Linear {
  int32_t n_input = 0, n_registers = 0;
  std::vector<Instruction> code;
  std::vector<int32_t> output;  // registers
};
Linear linearize(const Function& f);

// Algorithmic operation counts (cuda.rs:229-254 categories) over the linear program.
struct OpCounts {
  int64_t alu = 0;   // add, sub, mul, sq, min, max, compare
  int64_t div = 0, sqrt = 0, exp = 0, ln = 0;
  int64_t neg = 0, select = 0, constants = 0;
  // F_alg with the frozen weights: ALU 1, exp 20, ln 20, div 8, sqrt 8.
  int64_t weighted() const { return alu + 20 * (exp + ln) + 8 * (div + sqrt); }
};
OpCounts count(const Linear& l);

// Peak number of values live at once in program order (inputs and instruction results from definition to last
// use; constants excluded, since a compiler rematerializes them): the register demand of a one-lane-per-state
// kernel before any rescheduling.
int32_t peak_live(const Linear& l);

// A program split at its dependence on the inputs from `first_varying` on: `shared` computes from the inputs before
// it (its own inputs, the same registers) the B values that the varying part reads (its outputs); `varying` takes
// the program's inputs followed by those B values (registers n_input..n_input+B-1) and computes the program's
// outputs. Constants the varying part reads are emitted in it again. Evaluating `shared` once and `varying` for many
// inputs that differ only from `first_varying` on gives, instruction for instruction, the program's results.
struct Split {
  Linear shared, varying;
};
Split split_at(const Linear& l, int32_t first_varying);

}  // namespace combustion::ast
