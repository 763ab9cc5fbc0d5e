// C ABI of the MI355X HEOM integrator.  Declarations and the reference
// interfaces each entry point replaces: include/phi_hip.h.
#include "../../include/phi_hip.h"

#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "capi_internal.h"
#include "engine.h"
#include "mlp.h"
#include "integrator.h"
#include "phi_parameters.h"

struct phi_params {
  phi::PhiParameters p;
};
struct phi_integrator {
  phi::HierarchyIntegrator* it;
  phi::PhiParameters* p;
};
struct phi_mlp {
  phi::Mlp* m;
};

namespace phi {
namespace capi {

std::string& LastError() {
  thread_local std::string error;
  return error;
}

}  // namespace capi
}  // namespace phi

namespace {

using phi::capi::Err;
using phi::capi::Guarded;
using phi::capi::Ok;

phi::RunMethod MethodFromWord(const std::string& m, bool* known) {
  *known = true;
  if (m == "rk4spectrum") return phi::RK4SPECTRUM;
  if (m == "rkf45spectrum") return phi::RKF45SPECTRUM;
  if (m == "rk4") return phi::RK4;
  if (m == "rkf45") return phi::RKF45;
  if (m == "print") return phi::PRINTHIERARCHY;
  if (m == "memory") return phi::PRINTMEMORY;
  if (m == "bicgstabl" || m == "bicgstab") return phi::BICGSTABL;
  if (m == "bicgstabu") return phi::BICGSTABU;
  *known = false;
  return phi::NONE;
}

double* DoubleField(phi::PhiParameters& p, const std::string& n) {
  if (n == "hbar") return &p.hbar;
  if (n == "boltzmann_constant") return &p.boltzmann_constant;
  if (n == "minimum_field_change") return &p.minimum_field_change;
  if (n == "temperature") return &p.temperature;
  if (n == "integration_time") return &p.integration_time;
  if (n == "timestep") return &p.timestep;
  if (n == "output_minimum_timestep") return &p.output_minimum_timestep;
  if (n == "start_time") return &p.start_time;
  if (n == "restart_output_time") return &p.restart_output_time;
  if (n == "rkf45_minimum_timestep") return &p.rkf45_minimum_timestep;
  if (n == "rkf45_tolerance") return &p.rkf45_tolerance;
  if (n == "bicgstab_tolerance") return &p.bicgstab_tolerance;
  if (n == "filter_tolerance") return &p.filter_tolerance;
  return nullptr;
}

int* IntField(phi::PhiParameters& p, const std::string& n) {
  if (n == "verbose") return &p.verbose;
  if (n == "num_states") return &p.num_states;
  if (n == "hilbert_space_num_elements") return &p.hilbert_space_num_elements;
  if (n == "hilbert_space_size") return &p.hilbert_space_size;
  if (n == "num_bath_couplings") return &p.num_bath_couplings;
  if (n == "num_matsubara_terms") return &p.num_matsubara_terms;
  if (n == "hierarchy_truncation_level") return &p.hierarchy_truncation_level;
  if (n == "matrix_count") return &p.matrix_count;
  if (n == "bicgstab_ell") return &p.bicgstab_ell;
  if (n == "num_threads") return &p.num_threads;
  if (n == "bicgstab_max_iterations") return &p.bicgstab_max_iterations;
  return nullptr;
}

bool* BoolField(phi::PhiParameters& p, const std::string& n) {
  if (n == "is_truncated_system") return &p.is_truncated_system;
  if (n == "use_time_local_truncation") return &p.use_time_local_truncation;
  if (n == "is_multiple_independent_baths") return &p.is_multiple_independent_baths;
  if (n == "is_spectrum_calculation") return &p.is_spectrum_calculation;
  if (n == "is_diagonal_bath_coupling") return &p.is_diagonal_bath_coupling;
  if (n == "is_full_bath_coupling") return &p.is_full_bath_coupling;
  if (n == "is_rho_normalized") return &p.is_rho_normalized;
  if (n == "is_restarted") return &p.is_restarted;
  if (n == "stupid_hierarchy_partition") return &p.stupid_hierarchy_partition;
  if (n == "are_parameters_valid") return &p.are_parameters_valid;
  return nullptr;
}

std::vector<double>* RealArray(phi::PhiParameters& p, const std::string& n) {
  if (n == "gamma") return &p.gamma;
  if (n == "lambda") return &p.lambda;
  if (n == "kappa") return &p.kappa;
  if (n == "bath_coupling_op") return &p.bath_coupling_op;
  if (n == "initial_hamiltonian_field") return &p.initial_hamiltonian_field;
  if (n == "final_hamiltonian_field") return &p.final_hamiltonian_field;
  return nullptr;
}

std::vector<std::complex<double>>* Matrix(phi::PhiParameters& p, const std::string& n) {
  if (n == "time_indepenent_hamiltonian") return &p.time_indepenent_hamiltonian;
  if (n == "initial_annealing_hamiltonian") return &p.initial_annealing_hamiltonian;
  if (n == "final_annealing_hamiltonian") return &p.final_annealing_hamiltonian;
  if (n == "initial_density_matrix") return &p.initial_density_matrix;
  if (n == "spectrum_initial_density_matrix") return &p.spectrum_initial_density_matrix;
  return nullptr;
}

std::string* StringField(phi::PhiParameters& p, const std::string& n) {
  if (n == "output_filename") return &p.output_filename;
  if (n == "restart_input_filename") return &p.restart_input_filename;
  if (n == "restart_output_filename") return &p.restart_output_filename;
  if (n == "restart_backup_filename") return &p.restart_backup_filename;
  if (n == "full_precision_output") return &p.full_precision_output;
  return nullptr;
}

}  // namespace

extern "C" {

const char* phi_hip_last_error(void) { return phi::capi::LastError().c_str(); }
int phi_hip_version(void) { return 100; }

int phi_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int phi_hip_main(int argc, const char* const* argv) {
  return Guarded(__func__, [&]() -> int {
    phi::PhiParameters p;
    p.ReadCommandLine(argc, argv);
    phi::HierarchyIntegrator it(&p);
    it.Launch();
    if (!p.are_parameters_valid) return Err(PHI_HIP_EPARAMS, "invalid parameters");
    if (it.status()) return Err(it.status(), it.error());
    return Ok();
  });
}

int phi_hip_params_new(phi_params** out) {
  return Guarded(__func__, [&]() -> int {
    if (!out) return Err(PHI_HIP_EINVAL, "null output pointer");
    *out = new phi_params;
    return Ok();
  });
}

int phi_hip_params_read(const char* path, const char* method, int threads, phi_params** out) {
  return Guarded(__func__, [&]() -> int {
    if (!path || !out) return Err(PHI_HIP_EINVAL, "null argument");
    phi_params* h = new phi_params;
    *out = h;
    {
      std::ifstream probe(path);
      if (!probe.is_open()) {
        h->p.are_parameters_valid = false;
        return Err(PHI_HIP_EIO, std::string("cannot open ") + path);
      }
    }
    if (threads > 0) h->p.num_threads = threads;
    if (method) {
      bool known;
      h->p.run_method = MethodFromWord(method, &known);
    }
    h->p.ReadInputFile(path);
    if (!h->p.are_parameters_valid || h->p.run_method == phi::BROKENINPUT)
      return Err(PHI_HIP_EPARAMS, "invalid parameters");
    return Ok();
  });
}

void phi_hip_params_free(phi_params* p) { delete p; }

int phi_hip_params_get_double(const phi_params* h, const char* name, double* out) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name || !out) return Err(PHI_HIP_EINVAL, "null argument");
    double* f = DoubleField(const_cast<phi::PhiParameters&>(h->p), name);
    if (!f) return Err(PHI_HIP_EINVAL, std::string("unknown double field ") + name);
    *out = *f;
    return Ok();
  });
}

int phi_hip_params_set_double(phi_params* h, const char* name, double v) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name) return Err(PHI_HIP_EINVAL, "null argument");
    double* f = DoubleField(h->p, name);
    if (!f) return Err(PHI_HIP_EINVAL, std::string("unknown double field ") + name);
    *f = v;
    return Ok();
  });
}

int phi_hip_params_get_int(const phi_params* h, const char* name, int* out) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name || !out) return Err(PHI_HIP_EINVAL, "null argument");
    phi::PhiParameters& p = const_cast<phi::PhiParameters&>(h->p);
    const std::string n = name;
    if (int* f = IntField(p, n)) { *out = *f; return Ok(); }
    if (bool* b = BoolField(p, n)) { *out = *b ? 1 : 0; return Ok(); }
    if (n == "run_method") { *out = (int)p.run_method; return Ok(); }
    if (n == "bulk_precision") { *out = (int)p.bulk_precision; return Ok(); }
    return Err(PHI_HIP_EINVAL, "unknown int field " + n);
  });
}

int phi_hip_params_set_int(phi_params* h, const char* name, int v) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name) return Err(PHI_HIP_EINVAL, "null argument");
    const std::string n = name;
    if (int* f = IntField(h->p, n)) { *f = v; return Ok(); }
    if (bool* b = BoolField(h->p, n)) { *b = v != 0; return Ok(); }
    if (n == "run_method") { h->p.run_method = (phi::RunMethod)v; return Ok(); }
    if (n == "bulk_precision") {
      h->p.bulk_precision = v ? phi::kBulkFp32 : phi::kBulkFp64;
      return Ok();
    }
    return Err(PHI_HIP_EINVAL, "unknown int field " + n);
  });
}

int phi_hip_params_get_array(const phi_params* h, const char* name, double* out, size_t cap,
                             size_t* n) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name) return Err(PHI_HIP_EINVAL, "null argument");
    phi::PhiParameters& p = const_cast<phi::PhiParameters&>(h->p);
    if (std::string(name) == "multiple_independent_bath_indices") {
      const auto& v = p.multiple_independent_bath_indices;
      if (n) *n = v.size();
      if (out) {
        if (cap < v.size()) return Err(PHI_HIP_EINVAL, "buffer too small");
        for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
      }
      return Ok();
    }
    std::vector<double>* a = RealArray(p, name);
    if (!a) return Err(PHI_HIP_EINVAL, std::string("unknown array field ") + name);
    if (n) *n = a->size();
    if (out) {
      if (cap < a->size()) return Err(PHI_HIP_EINVAL, "buffer too small");
      std::memcpy(out, a->data(), a->size() * sizeof(double));
    }
    return Ok();
  });
}

int phi_hip_params_set_array(phi_params* h, const char* name, const double* v, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name || (n && !v)) return Err(PHI_HIP_EINVAL, "null argument");
    if (std::string(name) == "multiple_independent_bath_indices") {
      h->p.multiple_independent_bath_indices.assign(n, 0);
      for (size_t i = 0; i < n; ++i) h->p.multiple_independent_bath_indices[i] = (int)v[i];
      return Ok();
    }
    std::vector<double>* a = RealArray(h->p, name);
    if (!a) return Err(PHI_HIP_EINVAL, std::string("unknown array field ") + name);
    a->assign(v, v + n);
    return Ok();
  });
}

int phi_hip_params_get_matrix(const phi_params* h, const char* name, double* out, size_t cap,
                              size_t* n) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name) return Err(PHI_HIP_EINVAL, "null argument");
    auto* m = Matrix(const_cast<phi::PhiParameters&>(h->p), name);
    if (!m) return Err(PHI_HIP_EINVAL, std::string("unknown matrix field ") + name);
    if (n) *n = m->size();
    if (out) {
      if (cap < m->size()) return Err(PHI_HIP_EINVAL, "buffer too small");
      std::memcpy(out, (const void*)m->data(), m->size() * sizeof(std::complex<double>));
    }
    return Ok();
  });
}

int phi_hip_params_set_matrix(phi_params* h, const char* name, const double* v, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name || (n && !v)) return Err(PHI_HIP_EINVAL, "null argument");
    auto* m = Matrix(h->p, name);
    if (!m) return Err(PHI_HIP_EINVAL, std::string("unknown matrix field ") + name);
    m->resize(n);
    for (size_t i = 0; i < n; ++i) (*m)[i] = std::complex<double>(v[2 * i], v[2 * i + 1]);
    return Ok();
  });
}

int phi_hip_params_get_string(const phi_params* h, const char* name, char* out, size_t cap) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name || !out || cap == 0) return Err(PHI_HIP_EINVAL, "null argument");
    std::string* s = StringField(const_cast<phi::PhiParameters&>(h->p), name);
    if (!s) return Err(PHI_HIP_EINVAL, std::string("unknown string field ") + name);
    if (s->size() + 1 > cap) return Err(PHI_HIP_EINVAL, "buffer too small");
    std::memcpy(out, s->c_str(), s->size() + 1);
    return Ok();
  });
}

int phi_hip_params_set_string(phi_params* h, const char* name, const char* v) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !name || !v) return Err(PHI_HIP_EINVAL, "null argument");
    std::string* s = StringField(h->p, name);
    if (!s) return Err(PHI_HIP_EINVAL, std::string("unknown string field ") + name);
    *s = v;
    return Ok();
  });
}

int phi_hip_params_fix(phi_params* h, const char* method) {
  return Guarded(__func__, [&]() -> int {
    if (!h) return Err(PHI_HIP_EINVAL, "null argument");
    if (method) {
      bool known;
      h->p.run_method = MethodFromWord(method, &known);
      if (!known) return Err(PHI_HIP_EINVAL, std::string("unknown method ") + method);
    }
    h->p.FixIncompatibleCombinations();
    h->p.matrix_count = phi::GetHierarchyNodeCountEstimate(
        h->p.hierarchy_truncation_level, h->p.num_bath_couplings * h->p.num_matsubara_terms);
    return h->p.are_parameters_valid ? Ok() : Err(PHI_HIP_EPARAMS, "invalid parameters");
  });
}

int phi_hip_params_header(const phi_params* h, int num_nodes, char* out, size_t cap, size_t* len) {
  return Guarded(__func__, [&]() -> int {
    if (!h) return Err(PHI_HIP_EINVAL, "null argument");
    std::ostringstream o;
    h->p.Write(&o, num_nodes);
    o << "#Total Number of matrices:" << num_nodes << "\n";
    o << phi::MemoryTable(h->p, num_nodes);
    o << "#-------------\n";
    const std::string s = o.str();
    if (len) *len = s.size();
    if (out) {
      if (cap < s.size() + 1) return Err(PHI_HIP_EINVAL, "buffer too small");
      std::memcpy(out, s.c_str(), s.size() + 1);
    }
    return Ok();
  });
}

int phi_hip_validate(const char* kind, const char* text, int* valid) {
  return Guarded(__func__, [&]() -> int {
    if (!kind || !text || !valid) return Err(PHI_HIP_EINVAL, "null argument");
    const std::string k = kind, t = text;
    bool ok = true, v;
    if (k == "int") v = phi::prm::ValidInteger(t, &ok);
    else if (k == "float") v = phi::prm::ValidFloat(t, &ok);
    else if (k == "bool") v = phi::prm::ValidBool(t, &ok);
    else if (k == "reals") v = phi::prm::ValidRealList(t, &ok);
    else if (k == "complexes") v = phi::prm::ValidComplexList(t, &ok);
    else return Err(PHI_HIP_EINVAL, "unknown kind " + k);
    *valid = (v && ok) ? 1 : 0;
    return Ok();
  });
}

int phi_hip_integrator_create(phi_params* h, phi_integrator** out) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !out) return Err(PHI_HIP_EINVAL, "null argument");
    phi_integrator* it = new phi_integrator;
    it->p = &h->p;
    it->it = new phi::HierarchyIntegrator(&h->p);
    *out = it;
    return Ok();
  });
}

int phi_hip_integrator_launch(phi_integrator* it) {
  return Guarded(__func__, [&]() -> int {
    if (!it) return Err(PHI_HIP_EINVAL, "null argument");
    it->it->Launch();
    if (it->it->status()) return Err(it->it->status(), it->it->error());
    return Ok();
  });
}

int phi_hip_integrator_density_matrix(phi_integrator* it, double* out, size_t cap) {
  return Guarded(__func__, [&]() -> int {
    if (!it || !out) return Err(PHI_HIP_EINVAL, "null argument");
    const size_t dim = (size_t)it->it->density_matrix_dim();
    const size_t ne = dim * dim;
    if (cap < ne) return Err(PHI_HIP_EINVAL, "buffer too small");
    std::complex<double>* d = it->it->density_matrix();
    std::memcpy(out, (const void*)d, ne * sizeof(std::complex<double>));
    delete[] d;
    return Ok();
  });
}

int phi_hip_integrator_state(phi_integrator* it, double* out, size_t cap) {
  return Guarded(__func__, [&]() -> int {
    if (!it || !out) return Err(PHI_HIP_EINVAL, "null argument");
    std::vector<std::complex<double>> s;
    int rc = it->it->GetState(&s);
    if (rc) return Err(rc, "no device state (Launch not run or failed)");
    if (cap < s.size()) return Err(PHI_HIP_EINVAL, "buffer too small");
    std::memcpy(out, (const void*)s.data(), s.size() * sizeof(std::complex<double>));
    return Ok();
  });
}

int phi_hip_integrator_info(phi_integrator* it, double* out, size_t cap) {
  return Guarded(__func__, [&]() -> int {
    if (!it || !out || cap < 8) return Err(PHI_HIP_EINVAL, "bad argument");
    const phi::RunInfo& r = it->it->info();
    out[0] = (double)r.steps;
    out[1] = (double)r.accepted;
    out[2] = r.final_time;
    out[3] = r.final_dt;
    out[4] = r.seconds;
    out[5] = r.bicgstab_iterations;
    out[6] = r.bicgstab_residual;
    out[7] = r.num_nodes;
    return Ok();
  });
}

void phi_hip_integrator_destroy(phi_integrator* it) {
  try {
    if (!it) return;
    delete it->it;
    delete it;
  } catch (...) {
  }
}

int phi_hip_create(const phi_params* h, int precision, phi_ctx** out) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !out) return Err(PHI_HIP_EINVAL, "null argument");
    phi::PhiParameters p = h->p;
    p.bulk_precision = precision ? phi::kBulkFp32 : phi::kBulkFp64;
    phi::Engine* e = nullptr;
    std::string why;
    int rc = phi::Engine::Create(p, &e, &why);
    if (rc) return Err(rc, why);
    // Node 0 starts at rho0 (InitializeDensityMatrix), all other ADOs at 0.
    const size_t ne = (size_t)p.num_states * p.num_states;
    if (p.initial_density_matrix.size() == ne) {
      std::vector<std::complex<double>> s(e->state_elems(), 0);
      for (size_t i = 0; i < ne; ++i) s[i] = p.initial_density_matrix[i];
      rc = e->SetState(s.data());
      if (rc) {
        std::string m = e->last_error();
        delete e;
        return Err(rc, m);
      }
    }
    *out = new phi_ctx{e};
    return Ok();
  });
}

int phi_hip_ensemble_create(const phi_params* h, int members, const double* H, const int* depths,
                            int precision, phi_ctx** out) {
  return Guarded(__func__, [&]() -> int {
    if (!h || !H || !depths || !out || members < 1) return Err(PHI_HIP_EINVAL, "bad argument");
    phi::PhiParameters p = h->p;
    p.bulk_precision = precision ? phi::kBulkFp32 : phi::kBulkFp64;
    phi::Engine* e = nullptr;
    std::string why;
    int rc = phi::Engine::CreateEnsemble(p, members, reinterpret_cast<const std::complex<double>*>(H),
                                         depths, &e, &why);
    if (rc) return Err(rc, why);
    *out = new phi_ctx{e};
    return Ok();
  });
}

int phi_hip_ensemble_members(const phi_ctx* c, int* counts, int cap) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    const int m = c->e->members();
    if (counts) {
      const auto& n = c->e->member_counts();
      for (int i = 0; i < m && i < cap; ++i) counts[i] = n[i];
    }
    return m;
  });
}

int phi_hip_ensemble_rho0(phi_ctx* c, double* out, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !out) return Err(PHI_HIP_EINVAL, "null argument");
    const size_t m = c->e->members() > 0 ? c->e->members() : 1;
    if (n < m * c->e->N() * c->e->N()) return Err(PHI_HIP_EINVAL, "buffer too small");
    int rc = c->e->GetMemberRho0(reinterpret_cast<std::complex<double>*>(out));
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

static int Rk4History(phi_ctx* c, long long nsteps, double dt, int every, double* out,
                      size_t n, int* count, double* ms) {
  if (!c || !out || !count || nsteps < 0 || every < 1) return Err(PHI_HIP_EINVAL, "bad argument");
  phi::Engine* e = c->e;
  const size_t ne = (size_t)e->N() * e->N();
  if ((nsteps + every - 1) / every > e->history_capacity())
    return Err(PHI_HIP_EINVAL, "more snapshots than the device history holds");
  std::vector<std::complex<double>> h;
  int rc = e->DrainHistory(&h);  // discard anything recorded before
  if (rc >= 0) rc = ms ? e->Rk4StepsTimed(nsteps, dt, ms, every) : e->Rk4Steps(nsteps, dt, 0, every);
  if (rc >= 0) rc = e->DrainHistory(&h);
  if (rc < 0) return Err(rc, e->last_error());
  const size_t k = h.size() / ne;
  if (k * ne > n) return Err(PHI_HIP_EINVAL, "history buffer too small");
  std::memcpy(out, h.data(), h.size() * sizeof(std::complex<double>));
  *count = (int)k;
  return Ok();
}

int phi_hip_rk4_history(phi_ctx* c, long long nsteps, double dt, int every, double* out,
                        size_t n, int* count) {
  return Guarded(__func__, [&]() -> int {
    return Rk4History(c, nsteps, dt, every, out, n, count, nullptr);
  });
}

int phi_hip_rk4_history_timed(phi_ctx* c, long long nsteps, double dt, int every, double* out,
                              size_t n, int* count, double* ms) {
  return Guarded(__func__, [&]() -> int {
    if (!ms) return Err(PHI_HIP_EINVAL, "null argument");
    return Rk4History(c, nsteps, dt, every, out, n, count, ms);
  });
}

void phi_hip_destroy(phi_ctx* c) {
  try {
    if (!c) return;
    delete c->e;
    delete c;
  } catch (...) {
  }
}

// ---- Closure network ----------------------------------------------------------

int phi_hip_mlp_create(int n_in, int n_hidden, int n_out, int max_batch, phi_mlp** out) {
  return Guarded(__func__, [&]() -> int {
    if (!out) return Err(PHI_HIP_EINVAL, "null argument");
    phi::Mlp* m = nullptr;
    std::string why;
    int rc = phi::Mlp::Create(n_in, n_hidden, n_out, max_batch, &m, &why);
    if (rc) return Err(rc, why);
    *out = new phi_mlp{m};
    return Ok();
  });
}

void phi_hip_mlp_destroy(phi_mlp* m) {
  try {
    if (!m) return;
    delete m->m;
    delete m;
  } catch (...) {
  }
}

long long phi_hip_mlp_param_count(const phi_mlp* m) {
  return m ? m->m->param_count() : PHI_HIP_EINVAL;
}

int phi_hip_mlp_set_params(phi_mlp* m, const float* flat, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!m || !flat) return Err(PHI_HIP_EINVAL, "null argument");
    if ((long long)n != m->m->param_count()) return Err(PHI_HIP_EINVAL, "parameter count mismatch");
    int rc = m->m->SetParams(flat);
    return rc ? Err(rc, m->m->last_error()) : Ok();
  });
}

int phi_hip_mlp_get_params(phi_mlp* m, float* flat, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!m || !flat) return Err(PHI_HIP_EINVAL, "null argument");
    if ((long long)n < m->m->param_count()) return Err(PHI_HIP_EINVAL, "buffer too small");
    int rc = m->m->GetParams(flat);
    return rc ? Err(rc, m->m->last_error()) : Ok();
  });
}

int phi_hip_mlp_forward(phi_mlp* m, const float* x, int batch, float* y) {
  return Guarded(__func__, [&]() -> int {
    if (!m || !x || !y) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = m->m->Forward(x, batch, y);
    return rc ? Err(rc, m->m->last_error()) : Ok();
  });
}

int phi_hip_mlp_gradients(phi_mlp* m, const float* x, const float* t, int batch, double* loss,
                          float* grads, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!m || !x || !t || !loss || !grads) return Err(PHI_HIP_EINVAL, "null argument");
    if ((long long)n < m->m->param_count()) return Err(PHI_HIP_EINVAL, "buffer too small");
    int rc = m->m->Gradients(x, t, batch, loss, grads);
    return rc ? Err(rc, m->m->last_error()) : Ok();
  });
}

int phi_hip_mlp_train(phi_mlp* m, const float* X, const float* T, long long samples, int batch,
                      int epochs, float lr, float lr_final, unsigned seed, double* epoch_loss) {
  return Guarded(__func__, [&]() -> int {
    if (!m || !X || !T) return Err(PHI_HIP_EINVAL, "null argument");
    if (!(lr > 0)) return Err(PHI_HIP_EINVAL, "learning rate must be positive");
    int rc = m->m->Train(X, T, samples, batch, epochs, lr, lr_final, seed, epoch_loss);
    return rc ? Err(rc, m->m->last_error()) : Ok();
  });
}

int phi_hip_closure_samples(phi_ctx* c, int shallow_depth, float* X, float* T,
                            long long cap_rows, long long* rows) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !X || !T || !rows) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->ClosureSamples(shallow_depth, X, T, cap_rows, rows);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_closure_attach(phi_ctx* c, phi_mlp* m) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !m) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->AttachClosure(m->m);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_closure_fallback(phi_ctx* c, double ymax, long long* fallen) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->ClosureFallback(ymax, fallen);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_closure_fused(phi_ctx* c, int on) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    c->e->set_closure_fused(on != 0);
    return Ok();
  });
}

int phi_hip_num_nodes(const phi_ctx* c) { return c ? c->e->num_nodes() : PHI_HIP_EINVAL; }
int phi_hip_num_modes(const phi_ctx* c) { return c ? c->e->tables().D : PHI_HIP_EINVAL; }

int phi_hip_get_index(const phi_ctx* c, int* out, size_t cap) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !out) return Err(PHI_HIP_EINVAL, "null argument");
    const auto& occ = c->e->tables().occ;
    if (cap < occ.size()) return Err(PHI_HIP_EINVAL, "buffer too small");
    for (size_t i = 0; i < occ.size(); ++i) out[i] = occ[i];
    return Ok();
  });
}

int phi_hip_set_state(phi_ctx* c, const double* s, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !s) return Err(PHI_HIP_EINVAL, "null argument");
    if (n != c->e->state_elems()) return Err(PHI_HIP_EINVAL, "state size mismatch");
    int rc = c->e->SetState(reinterpret_cast<const std::complex<double>*>(s));
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_get_state(phi_ctx* c, double* s, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !s) return Err(PHI_HIP_EINVAL, "null argument");
    if (n != c->e->state_elems()) return Err(PHI_HIP_EINVAL, "state size mismatch");
    int rc = c->e->GetState(reinterpret_cast<std::complex<double>*>(s));
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_get_rho0(phi_ctx* c, double* out, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !out) return Err(PHI_HIP_EINVAL, "null argument");
    if (n < (size_t)c->e->N() * c->e->N()) return Err(PHI_HIP_EINVAL, "buffer too small");
    int rc = c->e->GetNode0(reinterpret_cast<std::complex<double>*>(out));
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_rhs(phi_ctx* c, const double* in, double* out, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !in || !out) return Err(PHI_HIP_EINVAL, "null argument");
    if (n != c->e->state_elems()) return Err(PHI_HIP_EINVAL, "state size mismatch");
    int rc = c->e->Rhs(reinterpret_cast<const std::complex<double>*>(in),
                       reinterpret_cast<std::complex<double>*>(out));
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_rhs_nodes(phi_ctx* c, const int* nodes, int count, double* drho, size_t n) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !nodes || !drho || count < 1) return Err(PHI_HIP_EINVAL, "bad argument");
    const size_t NE = (size_t)c->e->N() * c->e->N();
    if (n != (size_t)count * NE) return Err(PHI_HIP_EINVAL, "output size is not count * N^2");
    int rc = c->e->RhsNodes(nodes, count, reinterpret_cast<std::complex<double>*>(drho));
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_rk4_steps(phi_ctx* c, long long nsteps, double dt) {
  return Guarded(__func__, [&]() -> int {
    if (!c || nsteps < 0) return Err(PHI_HIP_EINVAL, "bad argument");
    int rc = c->e->Rk4Steps(nsteps, dt, 0, 0);
    if (!rc) rc = c->e->Synchronize();
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_rk4_steps_timed(phi_ctx* c, long long nsteps, double dt, double* ms) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !ms || nsteps < 0) return Err(PHI_HIP_EINVAL, "bad argument");
    int rc = c->e->Rk4StepsTimed(nsteps, dt, ms);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_rk4_stage_times(phi_ctx* c, long long nsteps, double dt, double* ms) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !ms || nsteps < 0) return Err(PHI_HIP_EINVAL, "bad argument");
    int rc = c->e->Rk4StageTimes(nsteps, dt, ms);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_rkf45_attempt(phi_ctx* c, const double* coefs, double* err) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !coefs || !err) return Err(PHI_HIP_EINVAL, "null argument");
    phi::Rkf45Coefs cf{};
    int q = 0;
    for (int s = 0; s < 5; ++s)
      for (int i = 0; i <= s; ++i) cf.b[s][i] = coefs[q++];
    for (int i = 0; i < 4; ++i) cf.c4[i] = coefs[q++];
    for (int i = 0; i < 5; ++i) cf.c5[i] = coefs[q++];
    int rc = c->e->Rkf45Attempt(cf, err);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_rkf45_accept(phi_ctx* c) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->Rkf45Accept();
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_active_flags(phi_ctx* c, int* flags, int n, int* count) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !flags || !count) return Err(PHI_HIP_EINVAL, "null argument");
    if (n < c->e->num_nodes()) return Err(PHI_HIP_EINVAL, "flags buffer too small");
    const int rc = c->e->ActiveFlags(flags);
    if (rc < 0) return Err(rc, c->e->last_error());
    *count = rc;
    return Ok();
  });
}

int phi_hip_steady_state(phi_ctx* c, double tol, int max_iter, int* iters, double* resid) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->SteadyState(tol, max_iter, iters, resid, nullptr);
    if (rc == phi::kErrNotConverged) return Err(rc, "BiCGSTAB did not converge");
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_steady_state_l(phi_ctx* c, int ell, double tol, int max_iter, int* iters,
                           double* resid) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->SteadyStateL(ell, tol, max_iter, iters, resid, nullptr);
    if (rc == phi::kErrNotConverged) return Err(rc, "BiCGSTAB(ell) did not converge");
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_set_launch(phi_ctx* c, int gemm, int tile, int threads, int blocks) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->SetLaunch(gemm, tile, threads, blocks);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_set_stage_launch(phi_ctx* c, int stage, int threads, int blocks) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->SetStageLaunch(stage, threads, blocks);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_stage_launch(const phi_ctx* c, int* threads, int* blocks) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !threads || !blocks) return Err(PHI_HIP_EINVAL, "null argument");
    c->e->StageLaunch(threads, blocks);
    return Ok();
  });
}

int phi_hip_set_graphs(phi_ctx* c, int on) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    c->e->set_graphs(on != 0);
    return Ok();
  });
}

int phi_hip_get_graphs(phi_ctx* c, int* on) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !on) return Err(PHI_HIP_EINVAL, "null argument");
    *on = c->e->graphs() ? 1 : 0;
    return Ok();
  });
}

int phi_hip_launch_counts(const phi_ctx* c, long long* graph_steps, long long* eager_steps) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !graph_steps || !eager_steps) return Err(PHI_HIP_EINVAL, "null argument");
    *graph_steps = c->e->last_graph_steps();
    *eager_steps = c->e->last_eager_steps();
    return Ok();
  });
}

int phi_hip_autotune(phi_ctx* c, int steps, double dt, double* best_ms, int* shape) {
  return Guarded(__func__, [&]() -> int {
    if (!c || !best_ms || !shape || steps < 1 || !(dt > 0)) return Err(PHI_HIP_EINVAL, "bad argument");
    int rc = c->e->Autotune(steps, dt, best_ms, shape);
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

int phi_hip_set_rk4_form(phi_ctx* c, int form) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    if (form < 0 || form > 2) return Err(PHI_HIP_EINVAL, "RK4 form must be 0, 1 or 2");
    c->e->set_rk4_form(form);
    return Ok();
  });
}

int phi_hip_synchronize(phi_ctx* c) {
  return Guarded(__func__, [&]() -> int {
    if (!c) return Err(PHI_HIP_EINVAL, "null argument");
    int rc = c->e->Synchronize();
    return rc ? Err(rc, c->e->last_error()) : Ok();
  });
}

}  // extern "C"
