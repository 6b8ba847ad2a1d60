// The device-function export (the nekRK files of cuda.rs:296-337) as text, and as a module: the four exported files
// compiled together by hiprtc (dfloat = double, __NEKRK_DEVICE__ = a device function) behind a kernel that forms
// their inputs from SoA states the way a CFD code's caller would, so that the exported text itself is run and
// checked on the GPU.
#include <cstdio>
#include <sstream>

#include "builders.hpp"
#include "device.hpp"
#include "emit.hpp"
#include "export.hpp"
#include "forms.hpp"
#include "soa.hpp"

namespace combustion {

std::string export_text(const Mechanism& m, int which, int D, const std::string& prefix) {
  if (which < 0 || which > 3) throw ArgumentError("which is CB_EXPORT_RATES, _CONDUCTIVITY, _VISCOSITY or _DIFFUSIVITY");
  if (D < 2 || D > 8) throw ArgumentError("D out of range");
  if (which == 0)
    return emit_device_function(ast::linearize(rates_device_function(m)), {"pressure_R", "lnT", "T", "T2", "T3", "T4", "rcpT"},
                                "concentrations", prefix + "species_rates", DeviceOutput::Out);
  const TransportScales sc = transport_scales(m);
  const Polynomials p = fit_polynomials(m.species, sc.T0, D);
  std::vector<std::string> scalars;
  if (which == 1) scalars.push_back("sum_mole_proportions");
  if (which == 3) scalars = {"mean_molar_mass_VTN", "VT"};
  for (int k = 1; k < D; k++) scalars.push_back(k == 1 ? "lnT" : "lnT" + std::to_string(k));
  const char* names[] = {"", "conductivityNIVT", "viscosityIVT", "density_diffusivity"};
  const ast::Linear l = ast::linearize(transport_device_function(m.species, p, sc.T0, sc.viscosity, sc.conductivity, which));
  return emit_device_function(l, scalars, "mole_proportions", prefix + names[which], which == 3 ? DeviceOutput::Strided : DeviceOutput::Return);
}

namespace {

std::string hex(double x) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", x);
  return buf;
}

// The exported files followed by a kernel over SoA states: inputs [T, P, Y_0..Y_{K-1}], outputs [ω_0..ω_{Ka-1}, λ, μ,
// D_0..D_{K-1}] in SI units, the arguments formed as the reference's check does (cuda.rs:350-369): concentrations as
// the mole proportions, powers of ln(T/T0), √(T/T0) and the scales of transport.rs:180-182.
std::string export_module_text(const Mechanism& m, int D) {
  const size_t K = m.species.len(), Ka = m.active;
  const TransportScales sc = transport_scales(m);
  std::ostringstream s;
  s << "typedef double dfloat;\n#define __NEKRK_DEVICE__ static __device__\n";
  for (int which = 0; which < 4; which++) s << export_text(m, which, D, "nekrk_") << "\n";
  s << "#define K " << K << "\n#define KA " << Ka << "\n#define RGAS (1.380649e-23 * 6.02214076e23)\n";
  s << "static __constant__ double molar_mass[K] = {";
  for (size_t k = 0; k < K; k++) s << (k ? ", " : "") << hex(m.species.molar_mass[k]);
  s << "};\n";
  std::string lnT_args;
  for (int d = 1; d < D; d++) lnT_args += ", lnT[" + std::to_string(d) + "]";
  s << "extern \"C\" __global__ void __launch_bounds__(256) nekrk_chemistry(const double* const* __restrict__ in, double* const* "
       "__restrict__ out, unsigned n) {\n"
       "  const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  if (id >= n) return;\n"
       "  const double T = in[0][id], P = in[1][id];\n"
       "  const double pressure_R = P * (1.0 / RGAS), rcpT = 1.0 / T;\n"
       "  double x[K], sum_n = 0.0, nonbulk = 0.0;\n"
       "  for (int k = 0; k < K; k++) { x[k] = in[2 + k][id] * (1.0 / molar_mass[k]); sum_n += x[k]; }\n"
       "  const double total = pressure_R * rcpT, density = total / sum_n;\n"
       "  for (int k = 0; k < K - 1; k++) { x[k] = density * fmax(0.0, x[k]); nonbulk += x[k]; }\n"
       "  x[K - 1] = total - nonbulk;\n"
       "  double w[KA];\n"
       "  const double T2 = T * T;\n"
       "  nekrk_species_rates(pressure_R, log(T), T, T2, T * T2, T2 * T2, rcpT, x, w);\n"
       "  for (int k = 0; k < KA; k++) out[k][id] = w[k];\n"
       "  const double theta = T * "
    << hex(1.0 / sc.T0)
    << ", VT = sqrt(theta);\n"
       "  double lnT["
    << D
    << "];\n"
       "  lnT[0] = 1.0; lnT[1] = log(theta);\n"
       "  for (int d = 2; d < "
    << D
    << "; d++) lnT[d] = lnT[d - 1] * lnT[1];\n"
       "  double WN = 0.0;\n"
       "  for (int k = 0; k < K; k++) WN += molar_mass[k] * x[k];\n"
       "  out[KA][id] = "
    << hex(sc.conductivity) << " * VT * nekrk_conductivityNIVT(total" << lnT_args
    << ", x);\n"
       "  out[KA + 1][id] = "
    << hex(sc.viscosity) << " * VT * nekrk_viscosityIVT(" << lnT_args.substr(2)
    << ", x);\n"
       "  double d[K];\n"
       "  nekrk_density_diffusivity(0, WN * VT, VT"
    << lnT_args
    << ", x, d, 1);\n"
       "  for (int k = 0; k < K; k++) out[KA + 2 + k][id] = d[k] * ("
    << hex(sc.viscosity)
    << " / WN);\n"
       "}\n";
  return s.str();
}

class ExportModule : public SoAForm {
 public:
  ExportModule(const Mechanism& m, int D) : SoAForm((int)m.species.len() + 2, (int)(m.active + 2 + m.species.len())) {
    kernel_.load(device::compile(export_module_text(m, D), "nekrk_chemistry", hip_options(false)), "nekrk_chemistry");
  }

 protected:
  void launch(const double* const* in, double* const* out, size_t n, const Launch& how) override {
    hipStream_t stream = how.stream;
    unsigned n32 = (unsigned)n, block = 256;
    void* args[] = {&in, &out, &n32};
    device::check(hipModuleLaunchKernel(kernel_.function, (unsigned)((n + block - 1) / block), 1, 1, block, 1, 1, 0, stream, args, nullptr),
                  "launch nekrk_chemistry");
  }

 private:
  device::Kernel kernel_;
};

}  // namespace

void compile_export(const Mechanism& m, int D) { device::compile(export_module_text(m, D), "nekrk_chemistry", hip_options(false)); }

std::unique_ptr<Form> assemble_export(const Mechanism& m, int D) { return std::make_unique<ExportModule>(m, D); }

}  // namespace combustion
