// extern "C" boundary (include/combustion_amd.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>

#include "../../include/combustion_amd.h"
#include "builders.hpp"
#include "device.hpp"
#include "emit.hpp"
#include "export.hpp"
#include "host_jit.hpp"
#include "ignition.hpp"
#include "integrator.hpp"
#include "jacobian.hpp"
#include "noalloc.hpp"
#include "forms.hpp"
#include "mechanism.hpp"
#include "mlp.hpp"
#include "yaml.hpp"

using namespace combustion;

struct cb_mechanism {
  std::shared_ptr<const Mechanism> m;
};

struct cb_program {
  std::shared_ptr<const Mechanism> mechanism;
  ast::Function function;
  ast::Linear linear;
  ProgramKind kind = ProgramKind::Other;
  ChemistryDescription chemistry;  // valid when kind == Chemistry
};

struct cb_module {
  std::unique_ptr<Form> form;
};

struct cb_cpu_module {
  std::unique_ptr<HostProgram> program;
};

struct cb_integrator {
  std::unique_ptr<Integrator> integrator;
};

struct cb_mlp {
  std::unique_ptr<Mlp> net;
};

namespace {

thread_local std::string last_error;

// An error already reported by a nested entry point: its status and message pass through unchanged.
struct StatusError : std::runtime_error {
  int status;
  StatusError(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

template <class F>
int guard(F&& f) {
  try {
    f();
    return CB_OK;
  } catch (const StatusError& e) {
    last_error = e.what();
    return e.status;
  } catch (const yaml::Error& e) {
    last_error = std::string("parse: ") + e.what();
    return CB_ERR_PARSE;
  } catch (const Unsupported& e) {
    last_error = e.what();
    return CB_ERR_UNSUPPORTED;
  } catch (const ast::Error& e) {
    last_error = std::string("build: ") + e.what();
    return std::string(e.what()).find("not implemented") != std::string::npos ? CB_ERR_UNSUPPORTED : CB_ERR_BUILD;
  } catch (const CompileError& e) {
    last_error = e.what();
    return CB_ERR_COMPILE;
  } catch (const device::Error& e) {
    last_error = e.what();
    return std::string(e.what()).rfind("hiprtc", 0) == 0 ? CB_ERR_COMPILE : CB_ERR_DEVICE;
  } catch (const ArgumentError& e) {
    last_error = e.what();
    return CB_ERR_ARGUMENT;
  } catch (const std::exception& e) {
    last_error = e.what();
    return CB_ERR_BUILD;
  }
}

void require(bool ok, const char* what) {
  if (!ok) throw ArgumentError(what);
}

int write_text(const std::string& s, char* buffer, size_t capacity, size_t* needed) {
  if (needed) *needed = s.size() + 1;
  if (!buffer || capacity < s.size() + 1) {
    last_error = "buffer too small";
    return CB_ERR_SIZE;
  }
  std::memcpy(buffer, s.c_str(), s.size() + 1);
  return CB_OK;
}

std::string num(double x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x > 0 ? "Infinity" : "-Infinity";
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", x);
  return buf;
}

std::string mechanism_json(const Mechanism& m) {
  const Species& s = m.species;
  std::ostringstream o;
  o << "{\"active\":" << m.active << ",\"species\":[";
  for (size_t k = 0; k < s.len(); k++) {
    if (k) o << ",";
    o << "{\"name\":\"" << s.names[k] << "\",\"molar_mass\":" << num(s.molar_mass[k]) << ",\"composition\":[";
    for (int e = 0; e < n_elements; e++) o << (e ? "," : "") << s.composition[k][e];
    const NASA7& t = s.thermodynamics[k];
    o << "],\"temperature_split\":" << num(t.temperature_split) << ",\"pieces\":[";
    for (int p = 0; p < 2; p++) {
      o << (p ? ",[" : "[");
      for (int i = 0; i < 7; i++) o << (i ? "," : "") << num(t.pieces[p][i]);
      o << "]";
    }
    o << "],\"geometry\":" << (int)s.geometry[k] << ",\"diameter\":" << num(s.diameter[k]) << ",\"well_depth_J\":" << num(s.well_depth_J[k])
      << ",\"polarizability\":" << num(s.polarizability[k]) << ",\"permanent_dipole_moment\":" << num(s.permanent_dipole_moment[k])
      << ",\"rotational_relaxation\":" << num(s.rotational_relaxation[k]) << ",\"internal_degrees_of_freedom\":" << num(s.internal_degrees_of_freedom[k])
      << "}";
  }
  o << "],\"reactions\":[";
  static const char* models[] = {"elementary", "irreversible", "three-body", "pressure-modification", "troe", "sri"};
  for (size_t r = 0; r < m.reactions.size(); r++) {
    const Reaction& x = m.reactions[r];
    if (r) o << ",";
    o << "{\"equation\":\"" << x.equation << "\",\"model\":\"" << models[(int)x.model] << "\",\"reactants\":[";
    for (size_t k = 0; k < x.reactants.size(); k++) o << (k ? "," : "") << (int)x.reactants[k];
    o << "],\"products\":[";
    for (size_t k = 0; k < x.products.size(); k++) o << (k ? "," : "") << (int)x.products[k];
    o << "],\"net\":[";
    for (size_t k = 0; k < x.net.size(); k++) o << (k ? "," : "") << (int)x.net[k];
    o << "],\"sum_net\":" << x.sum_net;
    auto rc = [&](const char* name, const RateConstant& c) {
      o << ",\"" << name << "\":[" << num(c.preexponential_factor) << "," << num(c.temperature_exponent) << "," << num(c.activation_temperature) << "]";
    };
    rc("rate_constant", x.rate_constant);
    rc("k0", x.k0);
    o << ",\"troe\":[" << num(x.troe.A) << "," << num(x.troe.T3) << "," << num(x.troe.T1) << "," << num(x.troe.T2) << "]";
    o << ",\"sri\":[" << num(x.sri.A) << "," << num(x.sri.B) << "," << num(x.sri.C) << "," << num(x.sri.D) << "," << num(x.sri.E) << "]";
    o << ",\"plog\":[";
    for (size_t i = 0; i < x.plog.size(); i++) {
      const RateConstant& k = x.plog[i].k;
      o << (i ? ",[" : "[") << num(x.plog[i].P) << "," << num(k.preexponential_factor) << "," << num(k.temperature_exponent) << ","
        << num(k.activation_temperature) << "]";
    }
    o << "]";
    o << ",\"efficiencies\":[";
    for (size_t k = 0; k < x.efficiencies.size(); k++) o << (k ? "," : "") << num(x.efficiencies[k]);
    o << "]}";
  }
  o << "]}";
  return o.str();
}

cb_program* make_program(std::shared_ptr<const Mechanism> m, ast::Function f, ProgramKind kind) {
  auto* p = new cb_program;
  p->mechanism = std::move(m);
  p->function = std::move(f);
  p->linear = ast::linearize(p->function);
  p->kind = kind;
  return p;
}

}  // namespace

extern "C" {

const char* cb_last_error(void) { return last_error.c_str(); }
const char* cb_version(void) { return "combustion_amd 0.1 (gfx950)"; }

static std::shared_ptr<const Mechanism> parse_mechanism(const std::string& text) {
  try {
    return std::make_shared<const Mechanism>(parse_yaml(text));
  } catch (const yaml::Error&) {
    throw;
  } catch (const std::exception& e) {
    throw yaml::Error(e.what());
  }
}

int cb_mechanism_parse_yaml(const char* text, cb_mechanism** out) {
  return guard([&] {
    require(text && out, "null argument");
    *out = new cb_mechanism{parse_mechanism(text)};
  });
}

int cb_mechanism_load_yaml(const char* path, cb_mechanism** out) {
  return guard([&] {
    require(path && out, "null argument");
    std::ifstream f(path);
    if (!f) throw ArgumentError(std::string("cannot open ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    *out = new cb_mechanism{parse_mechanism(ss.str())};
  });
}

void cb_mechanism_free(cb_mechanism* m) { delete m; }

int cb_mechanism_sizes(const cb_mechanism* m, int32_t* species, int32_t* active, int32_t* reactions) {
  return guard([&] {
    require(m, "null mechanism");
    if (species) *species = (int32_t)m->m->species.len();
    if (active) *active = (int32_t)m->m->active;
    if (reactions) *reactions = (int32_t)m->m->reactions.size();
  });
}

int cb_mechanism_json(const cb_mechanism* m, char* buffer, size_t capacity, size_t* needed) {
  int status = guard([&] { require(m, "null mechanism"); });
  if (status) return status;
  return write_text(mechanism_json(*m->m), buffer, capacity, needed);
}

int cb_ignition_state_parse(const cb_mechanism* m, const char* text, double* temperature, double* pressure, double* amounts) {
  return guard([&] {
    require(m && text && temperature && pressure && amounts, "null argument");
    const IgnitionState st = ignition_state(*m->m, text);
    *temperature = st.temperature;
    *pressure = st.pressure;
    std::copy(st.amounts.begin(), st.amounts.end(), amounts);
  });
}

int cb_ignition_state_load(const cb_mechanism* m, const char* path, double* temperature, double* pressure, double* amounts) {
  return guard([&] {
    require(m && path, "null argument");
    std::ifstream f(path);
    if (!f) throw ArgumentError(std::string("cannot open ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const int status = cb_ignition_state_parse(m, ss.str().c_str(), temperature, pressure, amounts);
    if (status) throw StatusError(status, last_error);
  });
}

int cb_transport_fits(const cb_mechanism* m, int32_t D, double T0, double* conductivity, double* viscosity, double* diffusivity) {
  return guard([&] {
    require(m && conductivity && viscosity && diffusivity, "null argument");
    Polynomials p = fit_polynomials(m->m->species, T0, D);
    std::copy(p.conductivityIVT.begin(), p.conductivityIVT.end(), conductivity);
    std::copy(p.VviscosityIVVT.begin(), p.VviscosityIVVT.end(), viscosity);
    std::copy(p.diffusivityITVT.begin(), p.diffusivityITVT.end(), diffusivity);
  });
}

int cb_transport_species(const cb_mechanism* m, int32_t what, int32_t a, int32_t b, double T, double* out) {
  return guard([&] {
    require(m && out, "null argument");
    const Species& s = m->m->species;
    require(a >= 0 && (size_t)a < s.len() && b >= 0 && (size_t)b < s.len(), "species index out of range");
    switch (what) {
      case 0: *out = transport::viscosity(s, a, T); break;
      case 1: *out = transport::conductivity(s, a, T); break;
      case 2: *out = transport::diffusivity(s, a, b, T); break;
      case 3: *out = transport::omega22(s, a, b, T); break;
      case 4: *out = transport::omega11(s, a, b, T); break;
      default: throw ArgumentError("unknown property");
    }
  });
}

int cb_transport_scales(const cb_mechanism* m, double T0, double* viscosity, double* conductivity) {
  return guard([&] {
    require(m && viscosity && conductivity, "null argument");
    TransportScales s = transport_scales(*m->m, T0);
    *viscosity = s.viscosity;
    *conductivity = s.conductivity;
  });
}

int cb_build_rates(const cb_mechanism* m, cb_program** out) {
  return guard([&] {
    require(m && out, "null argument");
    *out = make_program(m->m, rates(*m->m), ProgramKind::Other);
  });
}

int cb_build_reactions_rates(const cb_mechanism* m, cb_program** out) {
  return guard([&] {
    require(m && out, "null argument");
    *out = make_program(m->m, reactions_rates(*m->m), ProgramKind::Other);
  });
}

int cb_build_transport(const cb_mechanism* m, int32_t D, double T0, double viscosity, double conductivity, cb_program** out) {
  return guard([&] {
    require(m && out, "null argument");
    Polynomials p = fit_polynomials(m->m->species, T0, D);
    *out = make_program(m->m, properties(m->m->species, p, T0, viscosity, conductivity), ProgramKind::Other);
  });
}

int cb_build_chemistry(const cb_mechanism* m, int32_t flags, int32_t D, double T0, double viscosity, double conductivity, cb_program** out) {
  return guard([&] {
    require(m && out, "null argument");
    require((flags & 7) > 0 && flags < 16, "flags must combine CB_RATES, CB_TRANSPORT, CB_RHS (and CB_LEAN)");
    ChemistryDescription d;
    d.options.rates = flags & CB_RATES;
    d.options.transport = flags & CB_TRANSPORT;
    d.options.rhs = flags & CB_RHS;
    d.options.lean_rates = d.options.lean_transport = flags & CB_LEAN;
    d.scales = transport_scales(*m->m, T0);
    if (viscosity > 0) d.scales.viscosity = viscosity;
    if (conductivity > 0) d.scales.conductivity = conductivity;
    if (d.options.transport) d.polynomials = std::make_shared<Polynomials>(fit_polynomials(m->m->species, T0, D));
    ast::Function f = chemistry(*m->m, d.polynomials.get(), d.scales, d.options);
    cb_program* p = make_program(m->m, std::move(f), ProgramKind::Chemistry);
    p->chemistry = d;
    *out = p;
  });
}

int cb_build_reactor(const cb_mechanism* m, int32_t lean, cb_program** out) {
  return guard([&] {
    require(m && out, "null argument");
    *out = make_program(m->m, reactor(*m->m, lean != 0), ProgramKind::Other);
  });
}

int cb_build_jacobian(const cb_program* p, int32_t n_wrt, const int32_t* wrt, int32_t n_rows, const int32_t* rows, cb_program** out) {
  return guard([&] {
    require(p && out && n_wrt >= 0 && n_rows >= 0, "bad argument");
    require((n_wrt == 0 || wrt) && (n_rows == 0 || rows), "null index array");
    ast::Linear l = ast::jacobian(p->linear, std::vector<int32_t>(wrt, wrt + n_wrt), std::vector<int32_t>(rows, rows + n_rows));
    auto* j = new cb_program;
    j->mechanism = p->mechanism;
    j->function.n_input = l.n_input;
    j->function.output.assign(l.output.size(), -1);  // the program lives in its linear form only
    j->linear = std::move(l);
    j->kind = ProgramKind::Other;
    *out = j;
  });
}

int cb_program_split(const cb_program* p, int32_t first_varying, cb_program** shared, cb_program** varying) {
  return guard([&] {
    require(p && shared && varying, "null argument");
    require(first_varying >= 0 && first_varying <= p->linear.n_input, "first_varying out of range");
    ast::Split parts = ast::split_at(p->linear, first_varying);
    auto wrap = [&](ast::Linear l) {
      auto* q = new cb_program;
      q->mechanism = p->mechanism;
      q->function.n_input = l.n_input;
      q->function.output.assign(l.output.size(), -1);  // the program lives in its linear form only
      q->linear = std::move(l);
      q->kind = ProgramKind::Other;
      return q;
    };
    std::unique_ptr<cb_program> s(wrap(std::move(parts.shared)));
    *varying = wrap(std::move(parts.varying));
    *shared = s.release();
  });
}

void cb_program_free(cb_program* p) { delete p; }

int cb_program_sizes(const cb_program* p, int32_t* inputs, int32_t* outputs, int64_t* statements, int64_t* instructions, int32_t* registers) {
  return guard([&] {
    require(p, "null program");
    if (inputs) *inputs = p->function.n_input;
    if (outputs) *outputs = (int32_t)p->function.output.size();
    if (statements) *statements = (int64_t)p->function.statements.size();
    if (instructions) *instructions = (int64_t)p->linear.code.size();
    if (registers) *registers = p->linear.n_registers;
  });
}

int cb_program_op_counts(const cb_program* p, int64_t counts[8]) {
  return guard([&] {
    require(p && counts, "null argument");
    ast::OpCounts c = ast::count(p->linear);
    int64_t v[8] = {c.alu, c.div, c.sqrt, c.exp, c.ln, c.neg, c.select, c.constants};
    std::copy(v, v + 8, counts);
  });
}

int cb_program_linear(const cb_program* p, uint8_t* op, int32_t* dst, int32_t* a, int32_t* b, int32_t* c, double* k, int32_t* outputs) {
  return guard([&] {
    require(p && op && dst && a && b && c && k && outputs, "null argument");
    for (size_t i = 0; i < p->linear.code.size(); i++) {
      const ast::Instruction& x = p->linear.code[i];
      op[i] = (uint8_t)x.op;
      dst[i] = x.dst;
      a[i] = x.a;
      b[i] = x.b;
      c[i] = x.c;
      k[i] = x.k;
    }
    std::copy(p->linear.output.begin(), p->linear.output.end(), outputs);
  });
}

int cb_program_from_linear(int32_t n_inputs, int64_t n_instructions, const uint8_t* op, const int32_t* dst, const int32_t* a,
                           const int32_t* b, const int32_t* c, const double* k, int32_t n_outputs, const int32_t* outputs,
                           cb_program** out) {
  return guard([&] {
    require(out && n_inputs >= 0 && n_instructions >= 0 && n_outputs >= 0, "bad argument");
    require(n_instructions == 0 || (op && dst && a && b && c && k), "null instruction array");
    require(n_outputs == 0 || outputs, "null outputs");
    ast::Linear l;
    l.n_input = n_inputs;
    int32_t next = n_inputs;  // SSA: each instruction defines the next register
    auto defined = [&](int32_t r) { return r >= 0 && r < next; };
    for (int64_t i = 0; i < n_instructions; i++) {
      require(op[i] <= (uint8_t)ast::LOp::Select, "unknown op code");
      const auto o = (ast::LOp)op[i];
      require(dst[i] == next, "instructions must define registers n_inputs, n_inputs+1, ... in order");
      const int arity = o == ast::LOp::Const ? 0 : o == ast::LOp::Select ? 3
                        : (o == ast::LOp::Neg || o == ast::LOp::Sqrt || o == ast::LOp::Exp || o == ast::LOp::Ln || o == ast::LOp::Sq) ? 1 : 2;
      require(arity < 1 || defined(a[i]), "operand a is not a defined register");
      require(arity < 2 || defined(b[i]), "operand b is not a defined register");
      require(arity < 3 || defined(c[i]), "operand c is not a defined register");
      l.code.push_back({o, dst[i], a[i], b[i], c[i], k[i]});
      next++;
    }
    l.n_registers = next;
    for (int32_t i = 0; i < n_outputs; i++) {
      require(defined(outputs[i]), "output is not a defined register");
      l.output.push_back(outputs[i]);
    }
    auto* p = new cb_program;
    p->function.n_input = n_inputs;
    p->function.output.assign(n_outputs, -1);  // the program lives in its linear form only
    p->linear = std::move(l);
    p->kind = ProgramKind::Other;
    *out = p;
  });
}

int cb_export_device_function(const cb_mechanism* m, int32_t which, int32_t D, const char* prefix, char* buffer, size_t capacity,
                               size_t* needed) {
  std::string text;
  int status = guard([&] {
    require(m, "null mechanism");
    text = export_text(*m->m, which, D, prefix ? prefix : "");
  });
  if (status) return status;
  return write_text(text, buffer, capacity, needed);
}

int cb_hip_assemble_export(const cb_mechanism* m, int32_t D, cb_module** out) {
  return guard([&] {
    require(m, "null mechanism");
    require(D >= 2 && D <= 8, "D out of range");
    if (!out) return compile_export(*m->m, D);
    auto* mod = new cb_module;
    try {
      mod->form = assemble_export(*m->m, D);
    } catch (...) {
      delete mod;
      throw;
    }
    *out = mod;
  });
}

int cb_program_hip_source(const cb_program* p, int32_t form, uint32_t block_size, char* buffer, size_t capacity, size_t* needed) {
  std::string text;
  int status = guard([&] {
    require(p, "null program");
    ProgramView v{p->mechanism, &p->function, &p->linear, p->kind, &p->chemistry};
    text = form_source(v, form, block_size);
  });
  if (status) return status;
  return write_text(text, buffer, capacity, needed);
}

int cb_hip_compile(const cb_program* p, int32_t form, uint32_t block_size) {
  return guard([&] {
    require(p, "null program");
    ProgramView v{p->mechanism, &p->function, &p->linear, p->kind, &p->chemistry};
    compile_form(v, form, block_size);
  });
}

int cb_hip_assemble(const cb_program* p, int32_t form, uint32_t block_size, cb_module** out) {
  return guard([&] {
    require(p && out, "null argument");
    ProgramView v{p->mechanism, &p->function, &p->linear, p->kind, &p->chemistry};
    auto* m = new cb_module;
    try {
      m->form = assemble_form(v, form, block_size);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int cb_hip_call(cb_module* mod, const double* const* inputs, int32_t n_inputs, size_t n_states, double* const* outputs, int32_t n_outputs,
                int32_t pointers_are_device, void* stream, float* kernel_ms) {
  return guard([&] {
    require(mod && mod->form, "null module");
    require(n_inputs == mod->form->n_inputs() && n_outputs == mod->form->n_outputs(), "input/output count does not match the program");
    require(n_states == 0 || (inputs && outputs), "null array table");
    for (int32_t i = 0; i < n_inputs && n_states; i++) require(inputs[i], "null input array");
    for (int32_t i = 0; i < n_outputs && n_states; i++) require(outputs[i], "null output array");
    require(n_states < (size_t(1) << 31), "n_states must be below 2^31");
    mod->form->call(inputs, n_states, outputs, pointers_are_device != 0, (hipStream_t)stream, kernel_ms);
  });
}

int cb_hip_call_static(cb_module* mod, const double* const* inputs, int32_t n_inputs, size_t n_states, double* const* outputs,
                       int32_t n_outputs, void* stream) {
  return guard([&] {
    require(mod && mod->form, "null module");
    require(n_inputs == mod->form->n_inputs() && n_outputs == mod->form->n_outputs(), "input/output count does not match the program");
    require(n_states == 0 || (inputs && outputs), "null array table");
    for (int32_t i = 0; i < n_inputs && n_states; i++) require(inputs[i], "null input array");
    for (int32_t i = 0; i < n_outputs && n_states; i++) require(outputs[i], "null output array");
    require(n_states < (size_t(1) << 31), "n_states must be below 2^31");
    call_static(*mod->form, inputs, n_states, outputs, (hipStream_t)stream);
  });
}

void cb_module_free(cb_module* mod) { delete mod; }

int cb_cpu_assemble(const cb_program* p, cb_cpu_module** out) {
  return guard([&] {
    require(p && out, "null argument");
    auto* m = new cb_cpu_module;
    try {
      m->program = std::make_unique<HostProgram>(p->linear);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int cb_cpu_call(cb_cpu_module* mod, const double* const* inputs, int32_t n_inputs, size_t n_states, double* const* outputs, int32_t n_outputs,
                int32_t threads) {
  return guard([&] {
    require(mod && mod->program, "null module");
    require(n_inputs == mod->program->n_inputs() && n_outputs == mod->program->n_outputs(), "input/output count does not match the program");
    require(n_states == 0 || (inputs && outputs), "null array table");
    for (int32_t i = 0; i < n_inputs && n_states; i++) require(inputs[i], "null input array");
    for (int32_t i = 0; i < n_outputs && n_states; i++) require(outputs[i], "null output array");
    require(threads >= 1, "threads must be at least 1");
    mod->program->call(inputs, n_states, outputs, threads);
  });
}

void cb_cpu_module_free(cb_cpu_module* mod) { delete mod; }

int cb_integrator_create(const cb_mechanism* m, double rtol, double atol_T, double atol_Y, int32_t flags, cb_integrator** out) {
  return guard([&] {
    require(m, "null mechanism");
    if (!out) return compile_integrator(m->m, flags);
    auto* it = new cb_integrator;
    try {
      it->integrator = std::make_unique<Integrator>(m->m, rtol, atol_T, atol_Y, flags);
    } catch (...) {
      delete it;
      throw;
    }
    *out = it;
  });
}

int cb_integrator_advance(cb_integrator* it, size_t n_reactors, double* y, const double* P, double* t, double* h, int32_t* status,
                          uint32_t* steps, double t_end, int32_t max_steps, void* stream, int64_t* steps_taken) {
  return guard([&] {
    require(it && it->integrator, "null integrator");
    require(n_reactors == 0 || (y && P && t && h && status && steps), "null array");
    require(max_steps >= 0, "max_steps must be non-negative");
    const long taken = it->integrator->advance(y, P, t, h, status, steps, n_reactors, t_end, max_steps, (hipStream_t)stream);
    if (steps_taken) *steps_taken = taken;
  });
}

void cb_integrator_free(cb_integrator* it) { delete it; }

int cb_hip_assemble_surrogate(const cb_program* p, const cb_surrogate_weights* w, cb_module** out) {
  return guard([&] {
    require(p && w && out && w->params, "null argument");
    require(p->kind == ProgramKind::Chemistry, "the surrogate form needs a program built by cb_build_chemistry");
    const int NF = (int)p->mechanism->species.len() + 2;
    const MlpLayout lay(NF, w->width, w->depth, NF);
    SurrogateWeights sw;
    sw.width = w->width;
    sw.depth = w->depth;
    sw.params.assign(w->params, w->params + lay.n_params);
    auto copy = [&](std::vector<float>& v, const float* src) {
      require(src, "null normalisation array");
      v.assign(src, src + NF);
    };
    copy(sw.f_mean, w->feature_mean);
    copy(sw.f_rstd, w->feature_rstd);
    copy(sw.lo, w->lower);
    copy(sw.hi, w->upper);
    copy(sw.o_mean, w->output_mean);
    copy(sw.o_std, w->output_std);
    require(w->major && w->n_major >= 1 && w->n_major <= NF - 2, "n_major out of range");
    sw.major.assign(w->major, w->major + w->n_major);
    require(w->rates_form == 0 || w->rates_form == CB_FORM_FAST || w->rates_form == CB_FORM_MIXED,
            "rates_form is CB_FORM_FAST (or 0) or CB_FORM_MIXED");
    sw.rates_form = w->rates_form == CB_FORM_MIXED ? CB_FORM_MIXED : CB_FORM_FAST;
    ProgramView v{p->mechanism, &p->function, &p->linear, p->kind, &p->chemistry};
    auto* m = new cb_module;
    try {
      m->form = assemble_surrogate(v, sw);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int cb_module_fallback_count(cb_module* mod, void* stream, uint32_t* count) {
  return guard([&] {
    require(mod && mod->form && count, "null argument");
    *count = surrogate_fallback_count(*mod->form, (hipStream_t)stream);
  });
}

int cb_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, cb_mlp** out) {
  return guard([&] {
    require(out, "null argument");
    auto* m = new cb_mlp;
    try {
      m->net = std::make_unique<Mlp>(n_in, width, depth, n_out);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

void cb_mlp_free(cb_mlp* mlp) { delete mlp; }

int cb_mlp_sizes(const cb_mlp* mlp, int64_t* n_params, int32_t* in_pad, int32_t* out_pad, int64_t* layer_offsets) {
  return guard([&] {
    require(mlp && mlp->net, "null network");
    if (n_params) *n_params = mlp->net->n_params();
    if (in_pad) *in_pad = (int32_t)mlp->net->in_pad();
    if (out_pad) *out_pad = (int32_t)mlp->net->out_pad();
    if (layer_offsets)
      for (size_t l = 0; l < mlp->net->layers().size(); l++) {
        layer_offsets[2 * l] = mlp->net->layers()[l].w;
        layer_offsets[2 * l + 1] = mlp->net->layers()[l].b;
      }
  });
}

int cb_mlp_workspace_bytes(const cb_mlp* mlp, int64_t n, size_t* bytes) {
  return guard([&] {
    require(mlp && mlp->net && bytes && n >= 0, "bad argument");
    *bytes = mlp->net->workspace_bytes(n);
  });
}

int cb_mlp_forward(const cb_mlp* mlp, const float* params, const uint16_t* params_bf16, const float* x, int64_t n, float* y,
                   void* workspace, void* stream) {
  return guard([&] {
    require(mlp && mlp->net, "null network");
    require(n == 0 || (params && params_bf16 && x && y && workspace), "null argument");
    require(n < (int64_t(1) << 31), "n must be below 2^31");
    mlp->net->forward(params, params_bf16, x, n, y, workspace, (hipStream_t)stream);
  });
}

int cb_mlp_backward_mse(const cb_mlp* mlp, const uint16_t* params_bf16, const float* y, const float* target, int64_t n,
                        void* workspace, float* grad, void* stream) {
  return guard([&] {
    require(mlp && mlp->net, "null network");
    require(n == 0 || (params_bf16 && y && target && workspace && grad), "null argument");
    require(n < (int64_t(1) << 31), "n must be below 2^31");
    mlp->net->backward_mse(params_bf16, y, target, n, workspace, grad, (hipStream_t)stream);
  });
}

int cb_mlp_adam(const cb_mlp* mlp, float* params, uint16_t* params_bf16, const float* grad, float* m1, float* m2, float lr,
                float beta1, float beta2, float eps, int32_t step, void* stream) {
  return guard([&] {
    require(mlp && mlp->net && params && params_bf16 && grad && m1 && m2, "null argument");
    require(step >= 1, "Adam steps count from 1");
    mlp->net->adam(params, params_bf16, grad, m1, m2, lr, beta1, beta2, eps, step, (hipStream_t)stream);
  });
}

}  // extern "C"
