/* C ABI of the MI355X chemistry backend.
 *
 * The reference builds one straight-line program per mechanism (ast::Function) and hands it to a launcher:
 *   device::assemble(function, block_size) -> Fn(constants, SoA inputs) -> SoA outputs   (device.rs:25,75,98-100)
 * This header is that boundary for HIP on gfx950: the builders (yaml::parse, combustion::new, reaction::rates,
 * transport::properties) produce a cb_program, cb_hip_assemble compiles it with hiprtc into a cb_module, and
 * cb_hip_call runs it over N states in SoA layout. Plain pointers and sizes only.
 *
 * Conventions: every function returns CB_OK (0) or a negative status and never aborts; cb_last_error() then
 * holds a message for the calling thread. The caller owns every buffer (the reference returns freshly
 * allocated outputs, device.rs:34,72; here outputs are caller-allocated arrays of n_states doubles).
 * A cb_module is used by one host thread at a time; modules are independent.
 */
#ifndef COMBUSTION_AMD_H
#define COMBUSTION_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cb_mechanism cb_mechanism;
typedef struct cb_program cb_program;
typedef struct cb_module cb_module;
typedef struct cb_cpu_module cb_cpu_module;
typedef struct cb_integrator cb_integrator;

enum cb_status {
  CB_OK = 0,
  CB_ERR_ARGUMENT = -1,    /* null pointer, size mismatch (device.rs:31-32 asserts) */
  CB_ERR_PARSE = -2,       /* malformed mechanism file (yaml.rs panics) */
  CB_ERR_UNSUPPORTED = -3, /* e.g. SRI falloff (reaction.rs:194-203 unimplemented!) */
  CB_ERR_BUILD = -4,       /* builder contract violation (ast/lib.rs asserts, f32 overflow of a constant) */
  CB_ERR_COMPILE = -5,     /* hiprtc failure */
  CB_ERR_DEVICE = -6,      /* HIP runtime failure (device.rs anyhow::Result errors) */
  CB_ERR_SIZE = -7         /* caller buffer too small; *needed says how much */
};

/* Program forms for cb_hip_assemble. */
enum cb_form {
  CB_FORM_STRAIGHT = 0, /* the program emitted as straight-line HIP, one lane per state (cuda.rs:101-134) */
  CB_FORM_FP64 = 1,     /* chemistry programs: FP64 rates and transport sections as separate kernels (straight-line rates
                           and transport up to 64 species; the table-driven rates kernel and the tiled MFMA transport
                           kernel above) */
  CB_FORM_MIXED = 2,    /* tiled mixed precision: BF16-split MFMA contractions, FP32 exp/log, FP64 ΔG/RT */
  CB_FORM_FAST = 3      /* FP64 straight-line rates with FP32 cores for exp and ln, FP32 transport */
};
/* Modifier ORed into the form of a chemistry program (forms 1-3): the transport section takes the tiled MFMA
 * kernel (Wilke sums and binary-diffusion polynomials as v_mfma_{f64,f32}_16x16x4 contractions) at any species
 * count, instead of the straight-line transport kernel that forms 1 and 3 use up to 64 species. */
enum cb_form_flags { CB_FORM_TILED_TRANSPORT = 0x10,
/* Modifier for forms 1 and 3: the rates section takes the table-driven FP64 kernel (8 states x 8 reaction groups per
 * wave, per-state arrays in LDS) that those forms use above 64 species, instead of the straight-line rates kernel. */
                     CB_FORM_TILED_RATES = 0x20 };

/* Sections of the per-state program built by cb_build_chemistry. CB_LEAN selects the shapes that the
 * one-lane-per-state FP64 kernels compile (the same values to a few ulps per operation): rate-of-progress products
 * started from each reaction's rate constant, one-constant Wilke and binary-diffusion pair terms, and diffusion
 * reciprocals four pairs at a time. No reference counterpart: it exists so that the interpreter can check those
 * shapes against the oracle. */
enum cb_chemistry_flags { CB_RATES = 1, CB_TRANSPORT = 2, CB_RHS = 4, CB_LEAN = 8 };

const char* cb_last_error(void);
const char* cb_version(void);

/* yaml.rs:54 parse + lib.rs:127 new: Cantera-YAML (cm/s/mol units) to species and reaction tables. */
int cb_mechanism_load_yaml(const char* path, cb_mechanism** out);
int cb_mechanism_parse_yaml(const char* text, cb_mechanism** out);
void cb_mechanism_free(cb_mechanism* m);
/* K species, the first `active` of which react (lib.rs:130-136), and R reactions. */
int cb_mechanism_sizes(const cb_mechanism* m, int32_t* species, int32_t* active, int32_t* reactions);
/* JSON dump of the parsed tables in SI units (lib.rs:26-125 fields), for inspection and tests. */
int cb_mechanism_json(const cb_mechanism* m, char* buffer, size_t capacity, size_t* needed);

/* cantera.rs:13-29: the initial state of a `<mechanism>.ignition` file, a YAML sequence whose first entry holds
 * `pressure` [Pa], `temperature` [K] and `X`, a map from every species name to its mole proportion (>= 0, any
 * normalisation), and no `volume`. Outputs the reference's State: T, P and the amounts in V = 1 m³,
 * n_k = (P/R) V / T · x_k / Σx (amounts[K], the mechanism's species order). A missing species, a negative
 * proportion or a given volume (all assertions in the reference) return CB_ERR_PARSE. */
int cb_ignition_state_load(const cb_mechanism* m, const char* path, double* temperature, double* pressure, double* amounts);
int cb_ignition_state_parse(const cb_mechanism* m, const char* text, double* temperature, double* pressure, double* amounts);

/* transport.rs:125-136 Polynomials::<D>::new: fits in ln(T/T0), conductivity[K][D], viscosity[K][D],
 * diffusivity[K][K][D] (entries j<k fitted, j>k mirrored). */
int cb_transport_fits(const cb_mechanism* m, int32_t D, double T0, double* conductivity, double* viscosity, double* diffusivity);
/* transport.rs:97-116 single-species physics at T: what = 0 viscosity(a), 1 conductivity(a),
 * 2 diffusivity(a,b), 3 Ω*22(a,b), 4 Ω*11(a,b). */
int cb_transport_species(const cb_mechanism* m, int32_t what, int32_t a, int32_t b, double T, double* out);
/* Reference scales used by cb_build_chemistry (the role of cantera.rs:60-76). */
int cb_transport_scales(const cb_mechanism* m, double T0, double* viscosity, double* conductivity);

/* reaction.rs:356 rates: inputs [P/R, R/P, T, V, n_0..n_{K-2}], outputs [T dT/dt, dV/dt, ω_0..ω_{Ka-1}]. */
int cb_build_rates(const cb_mechanism* m, cb_program** out);
/* reaction.rs:329 reactions_rates: inputs [P/R, T, V, n_0..n_{K-2}], outputs one rate of progress per reaction. */
int cb_build_reactions_rates(const cb_mechanism* m, cb_program** out);
/* transport.rs:212 properties::<D>: inputs [Σx, T/T0, x_0..x_{K-2}], outputs [λ/λ0, μ/η0, ρD_0/η0..]. */
int cb_build_transport(const cb_mechanism* m, int32_t D, double T0, double viscosity, double conductivity, cb_program** out);
/* Per-state SoA program (the GPU boundary's shape, cuda.rs:210-226 family): inputs [T, P, Y_0..Y_{K-1}]
 * (K, Pa, mass fractions); outputs [ω_0..ω_{Ka-1}] if CB_RATES, [λ, μ, D_0..D_{K-1}] if CB_TRANSPORT,
 * [dT/dt] if CB_RHS, in that order. viscosity/conductivity <= 0 select cb_transport_scales. */
int cb_build_chemistry(const cb_mechanism* m, int32_t flags, int32_t D, double T0, double viscosity, double conductivity, cb_program** out);
/* The constant-pressure reactor's right-hand side in the variables a batch integrator steps (the role of the right-hand
 * side of cvode/lib.rs:17-47 and plot.rs:80-118): inputs [T, P, Y_0..Y_{K-1}], outputs [dT/dt, dY_0/dt..dY_{K-1}/dt],
 * dT/dt the energy balance of reaction.rs:375-380 and dY_k/dt = W_k ω_k / ρ with ρ = P / (R T Σ_j Y_j/W_j) (zero for
 * the species that do not react). lean != 0 takes the shapes of CB_LEAN. */
int cb_build_reactor(const cb_mechanism* m, int32_t lean, cb_program** out);
/* Analytic Jacobian by forward-mode differentiation of a program (no counterpart in the reference, whose integrator
 * takes CVODE's finite-difference Jacobians): a program with p's inputs whose outputs are p's outputs followed by
 * d output[rows[i]] / d input[wrt[j]], row-major (n_rows x n_wrt). min and max differentiate the operand they select
 * (the first on ties), a select its chosen operand. Assemble it like any other program. */
int cb_build_jacobian(const cb_program* p, int32_t n_wrt, const int32_t* wrt, int32_t n_rows, const int32_t* rows, cb_program** out);
/* A program split at its dependence on the inputs from first_varying on (no counterpart in the reference; the batched
 * integrator evaluates its difference-quotient copies this way): *shared computes, from the inputs before
 * first_varying, the B values the rest of the program reads (its outputs); *varying takes p's inputs followed by
 * those B values and computes p's outputs. shared once and varying per copy give p's results instruction for
 * instruction, so for many states that differ only in the varying inputs the shared part is evaluated once. */
int cb_program_split(const cb_program* p, int32_t first_varying, cb_program** shared, cb_program** varying);
void cb_program_free(cb_program* p);
int cb_program_sizes(const cb_program* p, int32_t* inputs, int32_t* outputs, int64_t* statements, int64_t* instructions, int32_t* registers);
/* cuda.rs:229-254 categories over the linear program: alu, div, sqrt, exp, ln, neg, select, constants. */
int cb_program_op_counts(const cb_program* p, int64_t counts[8]);
/* Linear SSA form of the program (ast/interpret.rs semantics, one instruction per node): op codes
 * 0 const,1 neg,2 min,3 max,4 add,5 sub,6 le,7 mul,8 div,9 sqrt,10 exp,11 ln,12 sq,13 select(a ? b : c).
 * Arrays have `instructions` entries; outputs has `outputs` entries (registers). */
int cb_program_linear(const cb_program* p, uint8_t* op, int32_t* dst, int32_t* a, int32_t* b, int32_t* c, double* k, int32_t* outputs);
/* The inverse of cb_program_linear: a program given in the linear form (same op codes and operand meaning),
 * for callers that build their own AST (the reference's ast::Function, linearised by its caller) and want the
 * straight-line HIP forms. Instruction i must define register n_inputs + i; operands refer to earlier registers. */
int cb_program_from_linear(int32_t n_inputs, int64_t n_instructions, const uint8_t* op, const int32_t* dst, const int32_t* a,
                           const int32_t* b, const int32_t* c, const double* k, int32_t n_outputs, const int32_t* outputs,
                           cb_program** out);
/* HIP text of a form (cuda.rs:101 compile counterpart). */
int cb_program_hip_source(const cb_program* p, int32_t form, uint32_t block_size, char* buffer, size_t capacity, size_t* needed);

/* Device-function export, the nekRK files of cuda.rs:140-350 (rates.c, conductivity.c, viscosity.c, diffusivity.c):
 * per-state `__NEKRK_DEVICE__` functions in `dfloat`, named <prefix>species_rates(pressure_R, lnT, T, T2, T3, T4,
 * rcpT, concentrations[], out[]) (out[k] = ω_k), <prefix>conductivityNIVT(sum_mole_proportions, lnT, lnT2..,
 * mole_proportions[]) and <prefix>viscosityIVT(lnT, lnT2.., mole_proportions[]) (returned), and
 * <prefix>density_diffusivity(id, mean_molar_mass_VTN, VT, lnT, lnT2.., mole_proportions[], out[], n_states)
 * (out[k*n_states+id]); lnT^m are powers of ln(T/T0), and the transport coefficients are fits of length D scaled by
 * cb_transport_scales (transport.rs:180-182). */
enum cb_export { CB_EXPORT_RATES = 0, CB_EXPORT_CONDUCTIVITY = 1, CB_EXPORT_VISCOSITY = 2, CB_EXPORT_DIFFUSIVITY = 3 };
int cb_export_device_function(const cb_mechanism* m, int32_t which, int32_t D, const char* prefix, char* buffer, size_t capacity,
                               size_t* needed);
/* The four exported files (prefix nekrk_, fits of length D) compiled together by hiprtc for gfx950 (dfloat = double,
 * __NEKRK_DEVICE__ = a device function) behind one kernel over SoA states, called with cb_hip_call: inputs
 * [T, P, Y_0..Y_{K-1}], outputs [ω_0..ω_{Ka-1}, λ, μ, D_0..D_{K-1}] in SI units, the functions' arguments formed as the
 * reference's check forms them (cuda.rs:350-369: mole proportions = concentrations, powers of ln(T/T0), √(T/T0)).
 * It runs the exported text itself on the device, as a CFD code including the files would. With out == NULL the
 * code object is only compiled into the cache (no GPU needed). */
int cb_hip_assemble_export(const cb_mechanism* m, int32_t D, cb_module** out);

/* device.rs:75 assemble: emit + hiprtc(gfx950) + module load. block_size 0 picks the form's default. */
int cb_hip_assemble(const cb_program* p, int32_t form, uint32_t block_size, cb_module** out);
/* Compiles a form into the code-object cache without touching a GPU. */
int cb_hip_compile(const cb_program* p, int32_t form, uint32_t block_size);
/* device.rs:59 call: inputs[n_inputs] and outputs[n_outputs] are SoA arrays of n_states doubles, host or
 * device memory (pointers_are_device). stream is a hipStream_t or NULL. With host pointers the call blocks;
 * with device pointers it is asynchronous unless kernel_ms is requested (then it waits and reports the
 * kernel time measured with HIP events on that stream). */
int cb_hip_call(cb_module* mod, const double* const* inputs, int32_t n_inputs, size_t n_states, double* const* outputs, int32_t n_outputs,
                int32_t pointers_are_device, void* stream, float* kernel_ms);
/* Allocation-free variant of cb_hip_call. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Device pointers
 * only. The arrays' pointer table, and a surrogate module's fallback counter and state list, live in one static
 * device array per GPU of 16 KiB + 256 B + 8 MiB, held until the process ends: at most 2048 arrays per module
 * (CB_ERR_UNSUPPORTED beyond) and at most 2^21 states per call on a surrogate module (CB_ERR_ARGUMENT beyond). The
 * module's kernels all run on `stream`, after small kernels that write the pointer table (nothing is copied from host
 * memory, so the call can be captured into a graph). Asynchronous. Every call on one GPU uses that GPU's one array:
 * calls that overlap on one GPU (different streams, several host threads) race; calls on one stream are ordered. */
int cb_hip_call_static(cb_module* mod, const double* const* inputs, int32_t n_inputs, size_t n_states, double* const* outputs,
                       int32_t n_outputs, void* stream);
void cb_module_free(cb_module* mod);

/* The CPU path: ir/lib.rs:225-263 compile + assemble (the reference's Cranelift JIT of one per-state function,
 * `extern fn(input: *const T, output: *mut T)`) and the loop of device.rs:28-48 around it. The program is emitted as
 * straight-line C in FP64 (libm exp/log, no FP contraction), compiled by the host compiler ($COMBUSTION_AMD_HOST_CC,
 * else ROCm's clang) at -O2 into a cached shared object and dlopen'ed. cb_cpu_call gathers each state's inputs
 * (SoA host arrays of n_states doubles) into one AoS buffer, calls the function and scatters its outputs, on
 * `threads` host threads over contiguous ranges of states. The benchmark's CPU baseline; the HIP entry points never
 * fall back to it. */
int cb_cpu_assemble(const cb_program* p, cb_cpu_module** out);
int cb_cpu_call(cb_cpu_module* mod, const double* const* inputs, int32_t n_inputs, size_t n_states, double* const* outputs, int32_t n_outputs,
                int32_t threads);
void cb_cpu_module_free(cb_cpu_module* mod);

/* Batched stiff integration of constant-pressure reactors on the GPU, the counterpart of the reference's CVODE
 * wrapper (cvode/lib.rs:17-47: CVodeCreate(CV_BDF), CVodeSStolerances(rtol, atol), dense linear solver, one step at
 * a time) and of the reactor loop of plot.rs:80-118, for many reactors at once: the stiffly accurate order-4
 * Rosenbrock method RODAS4 (Hairer & Wanner, with its embedded order-3 estimate; or, with CB_INTEGRATOR_ODE23S, the
 * L-stable (2,3) pair of ode23s, Shampine & Reichelt 1997; the error test in CVODE's weighted RMS norm) with the
 * Jacobian of cb_build_reactor's program (analytic, cb_build_jacobian, or difference quotients: flags) and
 * floor(64/(K+1)) reactors per wave for the LU factorisation and the solves (state size K+1 <= 64). Tolerances: rtol,
 * and atol_T (temperature) / atol_Y (mass fractions). With out == NULL the kernels are only compiled into the cache.
 * cb_integrator_advance moves n reactors toward t_end, in place, every array in device memory: y [K+1][n] (SoA rows
 * T, Y_0..Y_{K-1}), P [n] (constant), t [n] and h [n] (each reactor's time and next step size), status [n] (0
 * running; set to 1 at t_end, 2 when the step size collapsed; others are left alone), steps [n] (attempted steps).
 * It stops after max_steps steps of the batch or when no reactor runs (found by a poll every 8 steps; the steps
 * launched before the host reads that poll leave every reactor alone and are not counted in *taken), and blocks
 * until then. Parts of the batch run on streams of the integrator's own, after the work already on `stream` and
 * joined back into it. An integrator keeps device buffers sized for its last batch (allocated again, after a device
 * synchronization, when n, y, P or status change); calls on one integrator must not overlap. */
enum cb_integrator_flags {
  CB_JACOBIAN_AUTOMATIC = 0,   /* the analytic program up to 8 state rows (K <= 7), difference quotients above */
  CB_JACOBIAN_ANALYTIC = 1,    /* cb_build_jacobian's forward-mode program, f and J in one straight-line kernel */
  CB_JACOBIAN_DIFFERENCES = 2, /* difference quotients, as CVODE's dense solver forms them: every state copied K+2
                                  times (one copy per perturbed component, inc_j = sqrt(eps) max(|y_j|, atol_j/rtol)),
                                  one right-hand-side launch over the copies, the quotients formed by the LU kernel */
  CB_INTEGRATOR_ONE_PER_WAVE = 0x10, /* one reactor per Rosenbrock wave whatever its size (otherwise floor(64/(K+1))) */
  CB_INTEGRATOR_ODE23S = 0x20 /* ode23s's L-stable (2,3) pair (three stages) instead of RODAS4 */
};
int cb_integrator_create(const cb_mechanism* m, double rtol, double atol_T, double atol_Y, int32_t flags, cb_integrator** out);
int cb_integrator_advance(cb_integrator* it, size_t n_reactors, double* y, const double* P, double* t, double* h, int32_t* status,
                          uint32_t* steps, double t_end, int32_t max_steps, void* stream, int64_t* steps_taken);
void cb_integrator_free(cb_integrator* it);

/* Transport surrogate network (no counterpart in the reference, whose transport is always evaluated by the
 * transport.rs formulas; this is the learned replacement of that section). An MLP n_in -> width (SiLU) x depth ->
 * n_out with BF16 weights and FP32 accumulation on MFMA. Parameters are one flat FP32 array (FP32 master copy)
 * plus its BF16 copy, both device memory: per layer W [out][in] then b [out], with in/out padded to multiples
 * of 64 (cb_mlp_sizes). Activations are [n][in_pad] / [n][out_pad] FP32 row-major device arrays. */
/* Surrogate transport form of a chemistry program (inputs [T, P, Y], CB_TRANSPORT with or without CB_RATES):
 * ω by the FP32-core rates kernel (CB_FORM_FAST) or the MFMA mixed-precision one (CB_FORM_MIXED), and λ, μ, D_k by the
 * trained network (kernels/surrogate.hip), with the plain FP64 transport kernel (the MFMA-tiled one) evaluating the
 * states whose features fall outside the training box [lower, upper].
 * No counterpart in the reference. The network has n_in = n_out = K+2 (cb_mlp layout); features are
 * (T, ln P, X_0..X_{K-1}) (mole fractions) standardised by (mean, 1/std); outputs are [λ, μ, D_0..D_{K-1}] as
 * ln(property / baseline) standardised by (mean, std), the baselines costing O(K·n_major) per state: the
 * mixture rules with their pair sums over j restricted to the major species and the rest of the mole fraction
 * given to the bulk species (the last species; for the bulk itself, the first other major species). */
typedef struct cb_surrogate_weights {
  int32_t width, depth;
  const float* params;          /* cb_mlp_sizes(K+2, width, depth, K+2) n_params values */
  const float* feature_mean;    /* [K+2] */
  const float* feature_rstd;    /* [K+2] */
  const float* lower;           /* [K+2] training box, in feature units (T, ln P, X) */
  const float* upper;           /* [K+2] */
  const float* output_mean;     /* [K+2] */
  const float* output_std;      /* [K+2] */
  int32_t n_major;              /* 1..K, at least one major species other than the bulk */
  const int32_t* major;         /* [n_major] distinct species indices */
  int32_t rates_form;           /* the rates kernel beside the network: CB_FORM_FAST (or 0) for the FP32-core
                                   straight-line kernel, CB_FORM_MIXED for the MFMA mixed-precision kernel */
} cb_surrogate_weights;
int cb_hip_assemble_surrogate(const cb_program* p, const cb_surrogate_weights* w, cb_module** out);
/* States of the last cb_hip_call on a surrogate module that went to the plain transport kernel (waits for the
 * module's last stream). */
int cb_module_fallback_count(cb_module* mod, void* stream, uint32_t* count);

typedef struct cb_mlp cb_mlp;
int cb_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, cb_mlp** out);
void cb_mlp_free(cb_mlp* mlp);
/* n_params: flat parameter count; in_pad/out_pad: padded widths; layer_offsets[2 (depth+1)]: W and b offsets. */
int cb_mlp_sizes(const cb_mlp* mlp, int64_t* n_params, int32_t* in_pad, int32_t* out_pad, int64_t* layer_offsets);
int cb_mlp_workspace_bytes(const cb_mlp* mlp, int64_t n, size_t* bytes);
/* Forward pass (keeps activations in the workspace for cb_mlp_backward_mse). */
int cb_mlp_forward(const cb_mlp* mlp, const float* params, const uint16_t* params_bf16, const float* x, int64_t n, float* y,
                   void* workspace, void* stream);
/* Gradient of the mean squared error over the n x n_out valid outputs, written to grad (flat FP32). */
int cb_mlp_backward_mse(const cb_mlp* mlp, const uint16_t* params_bf16, const float* y, const float* target, int64_t n,
                        void* workspace, float* grad, void* stream);
/* One Adam step (step counts from 1) on the FP32 parameters; refreshes the BF16 copy. */
int cb_mlp_adam(const cb_mlp* mlp, float* params, uint16_t* params_bf16, const float* grad, float* m1, float* m2, float lr,
                float beta1, float beta2, float eps, int32_t step, void* stream);

#ifdef __cplusplus
}
#endif

#endif
