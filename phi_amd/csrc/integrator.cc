// HierarchyIntegrator for the MI355X HEOM integrator.  See integrator.h.
//
// Host control logic restates PHI's loops (hierarchy_integrator.cc):
//   RK4   2578-2939: while (time_now < end) { step; time_now += dt }, output
//         after step i when i % ceil(OutputMinimumTimestep / dt) == 0.
//   RKF45 1797-2573: Fehlberg weights pre-multiplied by the initial dt and
//         rescaled in place; accept when err < tol or dt/2 < mindt; grow by
//         1.1 when err < tol/10; the end clamp replaces the growth factor;
//         shrink by 0.5 on reject.
// The arithmetic of each step runs on the GPU (engine.cc / heom_kernels.hip).
#include "integrator.h"

#include "system_hamiltonian.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace phi {

namespace {

double Now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

std::string MemoryTable(const PhiParameters& p, int mc) {
  // 32-bit float sums in PHI's order; sizes: int 4, complex 16, double 8,
  // time_t 8.
  const long long n = p.num_states, m = p.num_bath_couplings, k = p.num_matsubara_terms;
  const long long T = p.num_threads, ne = n * n, hsne = p.hilbert_space_num_elements,
                  hs = p.hilbert_space_size;
  const long long C = 16, I = 4, F = 8, TT = 8;
  volatile float mem = 0;
  mem = mem + (float)(I * (m + 4 + m + m + m + m + 5) + C * (1 + m + m + m + m + ne));
  mem = mem * (float)(2 * mc);
  mem = mem + (float)(C * (3 + ne) + I * 6 + TT * 2);
  if (p.use_time_local_truncation) {
    mem = mem + (float)(C * ne * T);
    mem = mem + (float)(C * m * k * ne);
  }
  if (p.is_diagonal_bath_coupling) {
    mem = mem + (float)(C * m * ne * (k + 2) * T);
    mem = mem + (float)(C * m * hs);
  }
  if (p.is_full_bath_coupling) {
    mem = mem + (float)(C * m * ne * 3 * T);
    mem = mem + (float)(C * m * hsne);
  }
  mem = mem + (float)(C * hsne * T);
  if (p.is_truncated_system) {
    mem = mem + (float)(C * hs * T);
    mem = mem + (float)(C * hsne * T);
    mem = mem + (float)(I * 2 * T);
    mem = mem + (float)(C * ne * T);
  }
  volatile float matrices = (float)(C * ne * mc);
  matrices = matrices / 1024.0f;
  matrices = matrices / 1024.0f;
  volatile float rk4 = (float)(C * ne * mc * 4);
  rk4 = rk4 + mem;
  rk4 = rk4 / 1024.0f;
  rk4 = rk4 / 1024.0f;
  volatile float rkf = 0;
  rkf = rkf + (float)(C * ne * mc * 6);
  rkf = rkf + (float)(C * 31);
  rkf = rkf + (float)(F * 7 + I * 2);
  rkf = mem + rkf;
  rkf = rkf / 1024.0f;
  rkf = rkf / 1024.0f;
  volatile float mem2 = 0;
  mem2 = mem2 + (float)(I * (m + 4 + m + m + m + m + 5) + C * (1 + m + m + m + m + n));
  mem2 = mem2 * (float)(2 * mc);
  mem2 = mem2 + (float)(C * (3 + n) + I * 6 + TT * 2);
  volatile float srk4 = 0;
  srk4 = srk4 + (float)(C * (n + 1) * mc * 4);
  srk4 = srk4 + mem2;
  srk4 = srk4 / 1024.0f;
  srk4 = srk4 / 1024.0f;
  volatile float srkf = 0;
  srkf = srkf + (float)(C * 3 * n * mc * 6);
  srkf = srkf + (float)(C * 31);
  srkf = srkf + (float)(F * 7 + I * 2);
  srkf = srkf + mem2;
  srkf = srkf / 1024.0f;
  srkf = srkf / 1024.0f;
  volatile float ss = 0;
  ss = ss + (float)(I * (m + 4 + m + m + m + m + 5) + C * (1 + m + m + m + m));
  ss = ss + (float)(6 * (C * (n * n)));
  ss = ss * (float)mc;
  ss = ss + (float)(9 * C + 3 * I + 2 * F);
  ss = ss / 1024.0f;
  ss = ss / 1024.0f;
  std::ostringstream o;
  o << "#Memory to store matrices for calculations:\n"
    << "#hierarchy: " << (float)matrices << " Mb.\n"
    << "#rk4: " << (float)rk4 << " Mb.\n"
    << "#rkf45: " << (float)rkf << " Mb.\n"
    << "#rk4spectrum: " << (float)srk4 << " Mb.\n"
    << "#rkf45spectrum: " << (float)srkf << " Mb.\n"
    << "#steadystate: " << (float)ss << " Mb.\n";
  return o.str();
}

HierarchyIntegrator::HierarchyIntegrator(PhiParameters* p) : p_(p) {
  if (p_->verbose > 0) {
    std::cout << "PHI-compatible HEOM integrator on AMD Instinct MI355X (HIP, gfx950).\n"
              << "Integration of density matrix using hierarchy equations of motion.\n"
              << "\nPlease cite Strumpfer and Schulten, J. Chem. Theor. Comp. (2012)\n"
              << "in all publications reporting results obtained with PHI\n\n";
  }
}

HierarchyIntegrator::~HierarchyIntegrator() {
  delete engine_;
  delete sh_;
}

int HierarchyIntegrator::density_matrix_dim() const {
  return sh_ && sh_->truncated() ? sh_->full_size() : p_->num_states;
}

void HierarchyIntegrator::Fail(int code, const std::string& msg) {
  info_.status = code;
  error_ = msg;
  std::cerr << "ERROR: " << msg << "\n";
}

std::complex<double>* HierarchyIntegrator::density_matrix() {
  const int N = p_->num_states, F = density_matrix_dim();
  std::complex<double>* out = new std::complex<double>[(size_t)F * F];
  if (have_final_ && F != N) {
    sh_->UnTruncate(final_rho0_.data(), out);
  } else if (have_final_) {
    for (size_t i = 0; i < (size_t)N * N; ++i) out[i] = final_rho0_[i];
  } else {
    for (size_t i = 0; i < (size_t)F * F; ++i) out[i] = 0;
  }
  return out;
}

int HierarchyIntegrator::GetState(std::vector<std::complex<double>>* out) {
  if (!engine_) return kErrInvalidArgument;
  out->resize(engine_->state_elems());
  return engine_->GetState(out->data());
}

void HierarchyIntegrator::Launch() {
  info_ = RunInfo();
  const RunMethod rm = p_->run_method;
  if (!p_->are_parameters_valid || rm == BROKENINPUT) {
    if (p_->verbose > -1) std::cout << "Parameter error.\n";
    info_.status = kErrInvalidParameters;
    error_ = "invalid parameters";
  } else if (rm == RK4 || rm == RKF45 || rm == BICGSTAB || rm == BICGSTABL || rm == BICGSTABU) {
    int rc = CheckRun();
    if (!rc) rc = Initialize();
    if (!rc) rc = InitialState();
    if (!rc) {
      if (rm == RK4) rc = RunRk4();
      else if (rm == RKF45) rc = RunRkf45();
      else rc = RunSteadyState();
    }
    if (rc && info_.status == 0) info_.status = rc;
    // Launch() returns with the output complete on disk.
    if (out_.is_open()) out_.close();
    if (full_.is_open()) full_.close();
  } else if (rm == PRINTMEMORY) {
    std::cout << "Total number of matrices: " << p_->matrix_count << "\n";
    std::cout << "Number of elements per matrix: " << p_->num_states * p_->num_states << "\n";
    std::string t = MemoryTable(*p_, p_->matrix_count);
    // Console form: no leading '#', indented entries.
    std::istringstream in(t);
    std::string line;
    bool first = true;
    while (std::getline(in, line)) {
      if (first) { std::cout << line.substr(1) << "\n"; first = false; }
      else std::cout << "\t" << line.substr(1) << "\n";
    }
  } else if (rm == PRINTHIERARCHY) {
    HierarchyTables h;
    std::string msg = BuildHierarchy(*p_, &h);
    if (!msg.empty()) {
      Fail(kErrInvalidParameters, msg);
    } else {
      std::ostringstream ss;
      ss << "Hierarchy connections:\n";
      for (int l = h.Lh - 1; l >= 0; --l) {
        ss << "Level " << l << ":\n";
        for (int i = h.tier_start[l + 1] - 1; i >= h.tier_start[l]; --i) {
          ss << "\t" << i << " ->";
          for (int j = 0; j < h.D; ++j)
            if (h.prev[(size_t)i * h.D + j] >= 0) ss << " " << h.prev[(size_t)i * h.D + j];
          ss << "\n";
        }
      }
      std::cout << ss.str();
    }
  } else if (rm == RK4SPECTRUM || rm == RKF45SPECTRUM) {
    // PHI's spectrum mode has no defined result to reproduce: Get is not
    // virtual (hierarchy_updater.h:36), so its Spectrum* updaters run the
    // matrix DoUnitary / DoDephasing on N-entry vectors and write N^2
    // entries into N-entry buffers (hierarchy_integrator.cc:873-881).
    Fail(kErrUnsupported,
         "spectrum calculations are not supported (PHI's own spectrum mode writes N^2 "
         "entries into N-entry ADO buffers)");
  }
  if (p_->verbose > -2) std::cout << "Done.\n";
}

// Values the run loops need that PHI's reader guarantees but the C ABI's
// setters do not (a zero timestep or output period would loop forever),
// checked before anything is allocated.
int HierarchyIntegrator::CheckRun() {
  const PhiParameters& p = *p_;
  const RunMethod rm = p.run_method;
  std::string why;
  if (rm == RK4 || rm == RKF45) {
    if (!std::isfinite(p.integration_time) || !std::isfinite(p.start_time))
      why = "the start and end times must be finite";
    else if (!(p.timestep > 0) || !std::isfinite(p.timestep))
      why = "the timestep must be positive and finite";
    else if (!(p.output_minimum_timestep > 0))
      why = "OutputMinimumTimestep must be positive";
    else if (rm == RK4 && std::ceil(p.output_minimum_timestep / p.timestep) > 1e9)
      why = "OutputMinimumTimestep must be at most 1e9 timesteps";
    else if (rm == RK4 && (p.integration_time - p.start_time) / p.timestep > 1e15)
      why = "more than 1e15 RK4 steps requested";
    else if (rm == RKF45 && !(p.rkf45_tolerance > 0))
      why = "RKF45tolerance must be positive";
    else if (rm == RKF45 && !(p.rkf45_minimum_timestep > 0))
      why = "RKF45mindt must be positive";
  } else if (!(p.bicgstab_tolerance > 0)) {
    why = "the BiCGSTAB tolerance must be positive";
  }
  if (why.empty()) return kOk;
  Fail(kErrInvalidParameters, why);
  return info_.status;
}

int HierarchyIntegrator::Initialize() {
  const PhiParameters& p = *p_;
  std::string why;
  const PhiParameters* engine_params = &p;
  if (p.time_indepenent_hamiltonian.empty() || p.is_truncated_system) {
    // Annealing and/or truncated system: the engine integrates the
    // NumStates x NumStates Hamiltonian at the start of the run; the rest
    // is driven from here (hierarchy_integrator.cc:880-946, 1617-1667).
    sh_ = new SystemHamiltonian;
    why = sh_->Init(p);
    if (!why.empty()) {
      Fail(kErrInvalidParameters, why);
      return info_.status;
    }
    if (sh_->time_dependent() && p.run_method != RK4 && p.run_method != RKF45) {
      Fail(kErrUnsupported, "a time-dependent (annealing) Hamiltonian runs with RK4 or RKF45 only");
      return info_.status;
    }
    if (sh_->time_dependent() && p.is_restarted) {
      Fail(kErrUnsupported, "restarting a time-dependent (annealing) run is not supported");
      return info_.status;
    }
    ep_ = p;
    const int N = p.num_states;
    ep_.time_indepenent_hamiltonian.assign(sh_->H(), sh_->H() + (size_t)N * N);
    ep_.hilbert_space_size = N;
    ep_.hilbert_space_num_elements = N * N;
    ep_.is_truncated_system = false;
    engine_params = &ep_;
  }
  int rc = Engine::Create(*engine_params, &engine_, &why);
  if (rc) {
    Fail(rc, why);
    return rc;
  }
  // A truncated system's full coupling operators, in its eigenbasis.
  if (sh_ && sh_->BathOps()) {
    rc = engine_->SetBathOps(sh_->BathOps());
    if (!rc) rc = engine_->Synchronize();
    if (rc) {
      Fail(rc, engine_->last_error());
      return rc;
    }
  }
  info_.num_nodes = engine_->num_nodes();
  if (p.verbose > 0)
    std::cout << "\nTotal: " << engine_->num_nodes() << " matrices\n\n";
  WriteHeader();
  if (!p.full_precision_output.empty()) full_.open(p.full_precision_output.c_str());
  return kOk;
}

void HierarchyIntegrator::ActiveLine(std::ostream* os) {
  if (!engine_ || !engine_->filter_on() || p_->verbose <= 0) return;
  std::vector<int> flags(engine_->num_nodes());
  const int active = engine_->ActiveFlags(flags.data());
  if (active >= 0)
    *os << "#Total number of active matrices = " << active << "/" << engine_->num_nodes() << "\n";
}

void HierarchyIntegrator::WriteHeader() {
  if (p_->output_filename.empty()) return;
  out_.open(p_->output_filename.c_str());
  if (!out_.is_open()) return;
  p_->Write(&out_, engine_->num_nodes());
  out_ << "#Total Number of matrices:" << engine_->num_nodes() << std::endl;
  out_ << MemoryTable(*p_, engine_->num_nodes());
  out_ << "#-------------\n";
}

void HierarchyIntegrator::FullPrecisionLine(double t, const std::complex<double>* r) {
  if (!full_.is_open()) return;
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.17g", t);
  full_ << buf;
  const int ne = p_->num_states * p_->num_states;
  for (int i = 0; i < ne; ++i) {
    std::snprintf(buf, sizeof(buf), " %.17g %.17g", r[i].real(), r[i].imag());
    full_ << buf;
  }
  full_ << "\n";
}

int HierarchyIntegrator::InitialState() {
  const PhiParameters& p = *p_;
  const int N = p.num_states, NE = N * N;
  std::vector<std::complex<double>> state(engine_->state_elems(), 0);
  if (p.is_restarted) {
    double t = 0, dt = 0;
    int rc = ReadRestart(&t, &dt);
    if (rc) return rc;
    // A steady-state run writes dt = 0 (its file holds the solution, not a
    // point of a trajectory); time stepping cannot resume from it.
    if ((p.run_method == RK4 || p.run_method == RKF45) && !(dt > 0 && std::isfinite(dt))) {
      Fail(kErrInvalidParameters, "the restart file's timestep is not positive (a steady-state "
                                  "restart file cannot resume time stepping)");
      return info_.status;
    }
    if (!std::isfinite(t)) {
      Fail(kErrInvalidParameters, "the restart file's time is not finite");
      return info_.status;
    }
    p_->start_time = t;
    p_->timestep = dt;
    return kOk;
  }
  const int F = sh_ && sh_->truncated() ? sh_->full_size() : N;
  if ((int)p.initial_density_matrix.size() != F * F) {
    Fail(kErrUnsupported, "InitialDensityMatrix is required");
    return info_.status;
  }
  // ADOs stay Hermitian only from a Hermitian start; the integrator relies
  // on it (rho_n^dagger obeys the same equations).
  double scale = 0, asym = 0;
  for (int a = 0; a < F; ++a)
    for (int b = 0; b < F; ++b) {
      const std::complex<double> x = p.initial_density_matrix[a + (size_t)b * F];
      const std::complex<double> y = p.initial_density_matrix[b + (size_t)a * F];
      scale = std::max(scale, std::abs(x));
      asym = std::max(asym, std::abs(x - std::conj(y)));
    }
  if (asym > 1e-12 * std::max(1.0, scale)) {
    Fail(kErrInvalidParameters, "InitialDensityMatrix must be Hermitian");
    return info_.status;
  }
  // InitializeDensityMatrix (hierarchy_integrator.cc:1070-1091): node 0 gets
  // rho0 / scale with scale = 1 at n = 0; all other ADOs start at zero.  A
  // truncated system starts from rho0 in the eigenbasis (:1073-1076).
  if (F != N) sh_->Truncate(p.initial_density_matrix.data(), state.data());
  else
    for (int e = 0; e < NE; ++e) state[e] = p.initial_density_matrix[e];
  int rc = engine_->SetState(state.data());
  if (rc) {
    Fail(rc, engine_->last_error());
    return rc;
  }
  if (out_.is_open()) {
    out_ << p.start_time << " ";
    for (int e = 0; e < NE; ++e) out_ << state[e] << " ";
    out_ << "\n" << std::flush;
  }
  FullPrecisionLine(p.start_time, state.data());
  return kOk;
}

void HierarchyIntegrator::WriteRho0Line(const std::complex<double>* r, double t) {
  if (!out_.is_open()) return;
  out_ << t;
  const int ne = p_->num_states * p_->num_states;
  for (int j = 0; j < ne; ++j) out_ << " " << r[j];
  out_ << "\n";
}

void HierarchyIntegrator::WriteFinal(const std::complex<double>* r, double t) {
  final_rho0_.assign(r, r + (size_t)p_->num_states * p_->num_states);
  have_final_ = true;
  if (!out_.is_open()) return;
  out_ << std::setw(20) << std::setprecision(6) << t << std::setw(6) << std::setprecision(4);
  const int ne = p_->num_states * p_->num_states;
  for (int j = 0; j < ne; ++j) out_ << " " << r[j];
  out_ << "\n";
}

int HierarchyIntegrator::RunRk4() {
  const PhiParameters& p = *p_;
  const int NE = p.num_states * p.num_states;
  const double dt = p.timestep;
  const double end_time = p.integration_time;
  // time_now += dt must reach the end, and the output period must be at
  // least one step (PHI loops forever or divides by zero otherwise).
  if (!(dt > 0) || !std::isfinite(dt)) {
    Fail(kErrInvalidParameters, "the RK4 timestep must be positive and finite");
    return info_.status;
  }
  if (!std::isfinite(end_time) || !std::isfinite(p.start_time)) {
    Fail(kErrInvalidParameters, "the start and end times must be finite");
    return info_.status;
  }
  const double of = std::ceil(p.output_minimum_timestep / dt);
  if (!(of >= 1) || of > 1e9) {
    Fail(kErrInvalidParameters,
         "OutputMinimumTimestep must be positive and at most 1e9 timesteps");
    return info_.status;
  }
  const long long output_factor = (long long)of;
  if (end_time > p.start_time && (end_time - p.start_time) / dt > 1e15) {
    Fail(kErrInvalidParameters, "more than 1e15 RK4 steps requested");
    return info_.status;
  }
  if (p.verbose > -1) std::cout << "\tIntegrating using RK4.\n";
  // A long run first takes the fastest launch shape of the stage kernels
  // (the autotuner times about 50 steps and restores the state; results
  // change only in rounding).  Time-dependent systems and the explicit-k
  // paths (which the autotuner does not cover) keep theirs.
  constexpr double kAutotuneMinSteps = 10000;
  if (!(sh_ && sh_->time_dependent()) && end_time > p.start_time &&
      (end_time - p.start_time) / dt >= kAutotuneMinSteps) {
    double ms = 0;
    int shape[4];
    const int rc = engine_->Autotune(5, dt, &ms, shape);
    if (rc < 0 && rc != kErrUnsupported) {
      Fail(rc, engine_->last_error());
      return rc;
    }
  }
  const double t0 = Now();
  const std::time_t w0 = std::time(nullptr);
  std::time_t last_restart = w0;
  const bool restart = RestartOn();
  // Blocks of steps between host checks: the node-0 history never
  // overflows (2048 snapshots), and with restart output the wall clock is
  // read at least every 256 steps (PHI reads it after every step,
  // hierarchy_integrator.cc:2881-2895).
  long long chunk = output_factor * 2048;
  if (restart) chunk = std::min<long long>(chunk, 256);
  std::vector<std::complex<double>> hist;
  // Step times exactly as PHI accumulates them (F7: time_now += dt in
  // double until time_now >= end), generated one block at a time; t_before
  // is the time at the start of each step (a time-dependent H is evaluated
  // there).
  std::vector<double> times, t_before;
  double t = p.start_time;
  long long s = 0;
  while (t < end_time) {
    times.clear();
    t_before.clear();
    while (t < end_time && (long long)times.size() < chunk) {
      t_before.push_back(t);
      t += dt;
      times.push_back(t);
    }
    const long long n = (long long)times.size();
    int rc = sh_ && sh_->time_dependent()
                 ? engine_->Rk4StepsTD(n, t_before.data(), dt, s, (int)output_factor, sh_)
                 : engine_->Rk4Steps(n, dt, s, (int)output_factor);
    if (!rc) rc = engine_->DrainHistory(&hist);
    if (rc < 0) {
      Fail(rc, engine_->last_error());
      return rc;
    }
    // PHI's timing lines use whole seconds (time_t, difftime).
    const double el = std::difftime(std::time(nullptr), w0);
    long long k = 0;
    for (long long i = s; i < s + n; ++i) {
      if (i % output_factor != 0) continue;
      if (out_.is_open() && i % 100 == 0) {
        out_ << "#Time for 1st timestep: " << el / (double)(i + 1) << " s.\n"
             << "#Estimated time for 1 time unit: " << el / dt / 60 / 60 / (double)(i + 1)
             << " hrs.\n"
             << "#Total running time: " << 0 << " days.\n";
      }
      const std::complex<double>* r = hist.data() + (size_t)k * NE;
      WriteRho0Line(r, times[i - s]);
      FullPrecisionLine(times[i - s], r);
      ++k;
    }
    s += n;
    if (restart && t < end_time) {
      rc = MaybeWriteRestart(t, dt, &last_restart);
      if (rc) return rc;
    }
  }
  int rc = engine_->Synchronize();
  if (rc) {
    Fail(rc, engine_->last_error());
    return rc;
  }
  info_.seconds = Now() - t0;
  info_.steps = s;
  info_.final_time = t;
  info_.final_dt = dt;
  std::vector<std::complex<double>> r0(NE);
  rc = engine_->GetNode0(r0.data());
  if (rc) {
    Fail(rc, engine_->last_error());
    return rc;
  }
  if (p.verbose > -1) {
    std::ostringstream ss;
    ActiveLine(&ss);
    ss << "#Total number of integration steps = " << s << "\n"
       << "#Total integration time = " << std::difftime(std::time(nullptr), w0) / 60
       << " mins\n";
    std::cout << ss.str();
    if (out_.is_open()) out_ << ss.str();
  }
  WriteFinal(r0.data(), info_.final_time);
  FullPrecisionLine(info_.final_time, r0.data());
  if (restart) return WriteRestart(info_.final_time, dt);
  return kOk;
}

// A restart file when more than RestartOutputTimeSecs whole seconds of wall
// clock passed since the last one (PHI's difftime(now, last) > time,
// hierarchy_integrator.cc:2512-2526, 2881-2895).
int HierarchyIntegrator::MaybeWriteRestart(double t, double dt, std::time_t* last) {
  const std::time_t now = std::time(nullptr);
  if (!(std::difftime(now, *last) > p_->restart_output_time)) return kOk;
  const int rc = WriteRestart(t, dt);
  *last = std::time(nullptr);
  return rc;
}

int HierarchyIntegrator::RunRkf45() {
  const PhiParameters& p = *p_;
  const int NE = p.num_states * p.num_states;
  const double end_time = p.integration_time;
  const double tol = p.rkf45_tolerance, mindt = p.rkf45_minimum_timestep;
  double mydt = p.timestep, dt_f = p.timestep;
  double dtup = 1.1;
  const double dtdn = 0.5;
  if (!(mydt > 0)) {
    // time_now += dt would never reach the end.
    Fail(kErrInvalidParameters, "the RKF45 timestep must be positive");
    return info_.status;
  }
  // Step size before the end clamp shrank it (the restart file's dt).
  double free_dt = mydt;
  bool clamped = false;
  // Fehlberg weights (hierarchy_integrator.cc:1926-1975), each pre-multiplied
  // by the initial dt and then rescaled in place exactly as PHI does.
  double b[5][5] = {{2. / 9., 0, 0, 0, 0},
                    {1. / 12., 1. / 4., 0, 0, 0},
                    {69. / 128., -243. / 128., 135. / 64., 0, 0},
                    {-17. / 12., 27. / 4., -27. / 5, 16. / 15, 0},
                    {65. / 432., -5. / 16, 13. / 16., 4. / 27., 5. / 144.}};
  double c4[4] = {1. / 9, 9. / 20, 16. / 45, 1. / 12};
  double c5[5] = {47. / 450, 12. / 25, 32. / 225, 1. / 30, 6. / 25};
  // Stage times of a time-dependent Hamiltonian, scaled with the weights.
  double a[5] = {2. / 9, 1. / 3, 3. / 4, 1.0, 5. / 6};
  auto scale_all = [&](double f) {
    for (int s = 0; s < 5; ++s)
      for (int i = 0; i <= s; ++i) b[s][i] *= f;
    for (double& v : c4) v *= f;
    for (double& v : c5) v *= f;
    for (double& v : a) v *= f;
  };
  scale_all(dt_f);
  HamiltonianSource* src = sh_ && sh_->time_dependent() ? sh_ : nullptr;
  double time_now = p.start_time;
  if (!std::isfinite(end_time) || !std::isfinite(time_now)) {
    Fail(kErrInvalidParameters, "the start and end times must be finite");
    return info_.status;
  }
  // Dynamics lines of the output file.  PHI's writer compares an
  // uninitialised time_t `now` with next_write_time = start time
  // (hierarchy_integrator.cc:1808, 1840, 2500-2510); what it does in
  // practice, and what its example output shows (one_qubit_100.txt:49-53,
  // 447), is defined here: one line after the first accepted step; with
  // restart output on, `now` holds the wall clock of the previous accepted
  // step, so a further line follows whenever that wall clock passed the
  // previous line's by more than OutputMinimumTimestep seconds; plus the
  // t = 0 line and the final line.  The full-precision side file instead
  // takes a line after each accepted step at or past next_full_time
  // (simulated time), which then advances by OutputMinimumTimestep.
  const bool restart = RestartOn();
  const bool wall_clock = p.restart_output_time > 0;  // PHI then reads `now`
  bool first_line = true;
  std::time_t wall_prev = 0, wall_line = 0;
  double next_full_time = time_now;
  long long i = -1, accepted = 0, inext = 0;
  std::vector<std::complex<double>> r0(NE);
  const double t0 = Now();
  const std::time_t w0 = std::time(nullptr);
  std::time_t last_restart = w0;
  if (p.verbose > -1) std::cout << "\tIntegrating using adaptive timestepping RKF45.\n";
  while (time_now < end_time) {
    i += 1;
    Rkf45Coefs cf{};
    for (int s = 0; s < 5; ++s)
      for (int k = 0; k <= s; ++k) cf.b[s][k] = b[s][k];
    for (int k = 0; k < 4; ++k) cf.c4[k] = c4[k];
    for (int k = 0; k < 5; ++k) cf.c5[k] = c5[k];
    double max_diff = 0;
    int rc = engine_->Rkf45Attempt(cf, &max_diff, time_now, a, src);
    if (rc) {
      Fail(rc, engine_->last_error());
      return rc;
    }
    if (max_diff < tol || dt_f * dtdn < mindt) {
      rc = engine_->Rkf45Accept();
      if (rc) {
        Fail(rc, engine_->last_error());
        return rc;
      }
      time_now += mydt;
      accepted += 1;
      bool update_dt = max_diff < tol / 10;
      if (time_now + mydt > end_time) {
        if (!clamped) free_dt = update_dt ? mydt * dtup : mydt;
        clamped = true;
        update_dt = true;
        dtup = (end_time - time_now) / mydt;
      }
      if (update_dt) {
        mydt *= dtup;
        scale_all(dtup);
      }
      dt_f = mydt;
      if (out_.is_open() && i >= inext) {
        const double el = std::difftime(std::time(nullptr), w0);
        out_ << "#Info: Time for single timestep: " << el / (i + 1) << " s.\n"
             << "#Info: Estimated time for 1 time unit: "
             << (time_now > p.start_time ? el / (time_now - p.start_time) / 3600 : 0) << " hrs.\n"
             << "#Info: Total running time: "
             << (time_now > p.start_time ? el * end_time / (time_now - p.start_time) / 86400 : 0)
             << " days.\n";
        inext += 100;
      }
      bool line = first_line;
      if (!first_line && wall_clock &&
          std::difftime(wall_prev, wall_line) > p.output_minimum_timestep)
        line = true;
      const bool full_line = full_.is_open() && time_now >= next_full_time;
      if (line || full_line) {
        rc = engine_->GetNode0(r0.data());
        if (rc) {
          Fail(rc, engine_->last_error());
          return rc;
        }
      }
      if (line) {
        if (out_.is_open()) {
          out_ << std::setw(20) << std::setprecision(6) << time_now << std::setw(6)
               << std::setprecision(4);
          for (int j = 0; j < NE; ++j) out_ << " " << r0[j];
          out_ << std::endl;
        }
        wall_line = first_line ? std::time(nullptr) : wall_prev;
        first_line = false;
      }
      if (full_line) {
        FullPrecisionLine(time_now, r0.data());
        next_full_time = time_now + p.output_minimum_timestep;
      }
      if (wall_clock) wall_prev = std::time(nullptr);
      if (restart) {
        if (time_now < end_time) {
          // The step size the controller holds for the next attempt.
          rc = MaybeWriteRestart(time_now, clamped ? free_dt : dt_f, &last_restart);
          if (rc) return rc;
        }
      }
    } else {
      mydt *= dtdn;
      scale_all(dtdn);
      dt_f = mydt;
      if (dt_f < 1e-10) std::cout << "WARNING!!! dt < 1e-10\n";
    }
  }
  info_.seconds = Now() - t0;
  info_.steps = i + 1;
  info_.accepted = accepted;
  info_.final_time = time_now;
  info_.final_dt = dt_f;
  int rc = engine_->GetNode0(r0.data());
  if (rc) {
    Fail(rc, engine_->last_error());
    return rc;
  }
  if (p.verbose > -1) {
    std::ostringstream ss;
    ActiveLine(&ss);
    ss << "#Total number of integration steps = " << i + 1 << "\n"
       << "#Total number of accepted integration steps = " << accepted << "\n"
       << "#Total integration time = " << std::difftime(std::time(nullptr), w0) / 60
       << " mins.\n";
    std::cout << ss.str();
    if (out_.is_open()) out_ << ss.str();
  }
  WriteFinal(r0.data(), time_now);
  FullPrecisionLine(time_now, r0.data());
  // PHI writes dt_f, which the end clamp has shrunk to about (end - t) ~ 0,
  // so a run resumed from its file does not advance; the step size the
  // controller held before the clamp is written instead.
  if (restart) return WriteRestart(time_now, clamped ? free_dt : dt_f);
  return kOk;
}

int HierarchyIntegrator::RunSteadyState() {
  const PhiParameters& p = *p_;
  const int N = p.num_states, NE = N * N;
  int iters = 0;
  double resid = 0;
  std::vector<double> hist;
  const double t0 = Now();
  // PHI's bicgstab word with ell >= 1 selects BiCGSTAB(ell)
  // (hierarchy_integrator.cc:3112-3489), bicgstabu and ell = 0 plain BiCGSTAB.
  int rc = p.run_method == BICGSTABL && p.bicgstab_ell >= 1
               ? engine_->SteadyStateL(p.bicgstab_ell, p.bicgstab_tolerance,
                                       p.bicgstab_max_iterations, &iters, &resid, &hist)
               : engine_->SteadyState(p.bicgstab_tolerance, p.bicgstab_max_iterations, &iters,
                                      &resid, &hist);
  info_.seconds = Now() - t0;
  info_.bicgstab_iterations = iters;
  info_.bicgstab_residual = resid;
  if (rc && rc != kErrNotConverged) {
    Fail(rc, engine_->last_error());
    return rc;
  }
  std::vector<std::complex<double>> r0(NE);
  int rc2 = engine_->GetNode0(r0.data());
  if (rc2) {
    Fail(rc2, engine_->last_error());
    return rc2;
  }
  std::ostringstream ss;
  if (rc == kErrNotConverged)
    ss << "#Convergence failed! (steps = " << iters << ", diff = " << resid << ")\n";
  else
    ss << "Converged to diff = " << resid << " in " << iters << " steps.\n";
  std::cout << ss.str();
  if (out_.is_open()) {
    for (size_t k = 0; k < hist.size(); ++k) out_ << "#BiCGSTAB " << k << " " << hist[k] << "\n";
    out_ << ss.str();
    out_ << "#Info: Total running time: " << info_.seconds / 86400 << " days.\n";
    out_ << resid << " ";
    for (int j = 0; j < NE; ++j) out_ << r0[j] << " ";
    out_ << std::endl;
  }
  final_rho0_ = r0;
  have_final_ = true;
  FullPrecisionLine((double)iters, r0.data());
  if (RestartOn()) {
    int w = WriteRestart((double)iters, 0);
    if (w) return w;
  }
  if (rc == kErrNotConverged) {
    info_.status = kErrNotConverged;
    error_ = "BiCGSTAB did not converge";
  }
  return rc;
}

// Restart files: PHI's text layout (restart_helper.cc:30-55) - time, dt, then
// one line per node in id order holding its N*N column-major elements as
// "(re,im)" - written at 17 significant digits so that a resumed run is
// lossless; PHI's own reader (operator>>) reads it unchanged.
int HierarchyIntegrator::WriteRestart(double t, double dt) {
  std::vector<std::complex<double>> s;
  int rc = GetState(&s);
  if (rc) {
    Fail(rc, engine_->last_error());
    return rc;
  }
  // Written next to the target and renamed over it, so that a run killed
  // while writing leaves the previous file whole.
  const std::string name = p_->restart_output_filename;
  const std::string tmp = name + ".tmp";
  {
    std::ofstream o(tmp.c_str());
    if (!o.is_open()) {
      Fail(kErrIo, "cannot write restart file " + tmp);
      return kErrIo;
    }
    o << std::setprecision(17) << t << "\n" << dt << "\n";
    const size_t ne = (size_t)p_->num_states * p_->num_states;
    for (int n = 0; n < engine_->num_nodes(); ++n) {
      for (size_t i = 0; i < ne; ++i) o << (i ? " " : "") << s[n * ne + i];
      o << "\n";
    }
    // Time-local truncation: the saved matrices T_mk, one line each, bath
    // major (restart_helper.cc:43-52).  PHI saves them after accepted RKF45
    // steps only, so an RK4 run writes zeros here, as PHI's does.
    if (TimeLocal* tl = engine_->timelocal()) {
      const std::vector<std::complex<double>>& T = tl->saved();
      for (size_t mk = 0; mk < (size_t)tl->M() * tl->K(); ++mk) {
        for (size_t i = 0; i < ne; ++i) o << (i ? " " : "") << T[mk * ne + i];
        o << "\n";
      }
    }
    o.flush();
    if (!o.good()) {
      Fail(kErrIo, "cannot write restart file " + tmp);
      return kErrIo;
    }
  }
  if (std::rename(tmp.c_str(), name.c_str()) != 0) {
    Fail(kErrIo, "cannot replace restart file " + name);
    return kErrIo;
  }
  return kOk;
}

int HierarchyIntegrator::ReadRestart(double* t, double* dt) {
  std::ifstream in(p_->restart_input_filename.c_str());
  if (!in.is_open()) {
    Fail(kErrIo, "cannot read restart file " + p_->restart_input_filename);
    return kErrIo;
  }
  in >> *t >> *dt;
  std::vector<std::complex<double>> s(engine_->state_elems());
  for (auto& v : s) in >> v;
  // Time-local matrices: read as saved, then current = saved
  // (hierarchy_integrator.cc:1060-1068).
  TimeLocal* tl = engine_->timelocal();
  if (tl)
    for (auto& v : tl->saved()) in >> v;
  if (!in) {
    Fail(kErrIo, "restart file " + p_->restart_input_filename + " is truncated or malformed");
    return kErrIo;
  }
  int rc = engine_->SetState(s.data());
  if (!rc && tl) {
    tl->Restore();
    rc = engine_->TlUpload();
  }
  if (rc) {
    Fail(rc, engine_->last_error());
    return rc;
  }
  if (out_.is_open()) {
    out_ << *t << " ";
    for (int e = 0; e < p_->num_states * p_->num_states; ++e) out_ << s[e] << " ";
    out_ << "\n" << std::flush;
  }
  FullPrecisionLine(*t, s.data());
  return kOk;
}

}  // namespace phi
