// `.prm` reader and output header for the MI355X HEOM integrator.
//
// Behaviour follows PHI's PhiParameters (phi_parameters.cc:9-1147) and
// ParameterInput (parameter_input.cc): the same keys and aliases, the same
// line grammar (a line holding '#' is skipped; `Key=value` is a scalar;
// `Key:` starts a matrix whose rows follow one per line; a bare line continues
// the last matrix; a blank line ends it; a last line without '\n' is not read;
// unknown keys are ignored), the same value checks, the same defaults and the
// same output header.  The code is this project's own.
#include "phi_parameters.h"

#include <climits>
#include <cmath>

#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>

namespace phi {

namespace {

enum Range { kPositive, kNonNegative, kAnyRange };

bool IsDigit(char c) { return c >= '0' && c <= '9'; }

// Index of the first non-blank character; false when there is none (PHI
// reads past the end there and aborts; here the value is rejected).
bool SkipBlanks(const std::string& s, size_t* i) {
  *i = 0;
  while (*i < s.size() && s[*i] == ' ') ++*i;
  return *i < s.size();
}

}  // namespace

namespace prm {

// ^ *[0-9]+ *$ (parameter_input.cc:369-398).
bool ValidInteger(const std::string& s, bool* ok) {
  size_t i;
  *ok = SkipBlanks(s, &i);
  if (!*ok) return false;
  bool blank_seen = false;
  for (; i < s.size(); ++i) {
    const char c = s[i];
    if (c == ' ') { blank_seen = true; continue; }
    if (!IsDigit(c) || blank_seen) return false;
  }
  return true;
}

// One token over [0-9.eE+-] with at most one '.', never after the exponent;
// one exponent mark, not first; a sign only first or right after the
// exponent mark; only blanks after the token (parameter_input.cc:401-460).
bool ValidFloat(const std::string& s, bool* ok) {
  size_t i;
  *ok = SkipBlanks(s, &i);
  if (!*ok) return false;
  const size_t first = i;
  bool dot = false, expo = false, blank_seen = false;
  for (; i < s.size(); ++i) {
    const char c = s[i];
    if (c == ' ') { blank_seen = true; continue; }
    if (blank_seen) return false;
    const size_t pos = i - first;  // characters consumed before this one
    if (IsDigit(c)) continue;
    if (c == '.') {
      if (dot || expo) return false;
      dot = true;
    } else if (c == 'e' || c == 'E') {
      if (pos == 0 || expo) return false;
      expo = true;
    } else if (c == '-' || c == '+') {
      if (pos != 0 && s[i - 1] != 'e' && s[i - 1] != 'E') return false;
    } else {
      return false;
    }
  }
  return true;
}

// ^ *[01] *$ (parameter_input.cc:341-367).
bool ValidBool(const std::string& s, bool* ok) {
  size_t i;
  *ok = SkipBlanks(s, &i);
  if (!*ok) return false;
  if (s[i] != '0' && s[i] != '1') return false;
  for (++i; i < s.size(); ++i)
    if (s[i] != ' ') return false;
  return true;
}

// Comma-separated real numbers (parameter_input.cc:550-619).  Per field the
// float rules apply, except that blanks directly after a comma are free and
// a comma may follow blanks; no empty field, no trailing comma.
bool ValidRealList(const std::string& s, bool* ok) {
  size_t i;
  *ok = SkipBlanks(s, &i);
  if (!*ok) return false;
  bool dot = false, expo = false, after_comma = false, blank_seen = false;
  int pos = 0;  // position inside the current field
  while (i < s.size()) {
    const char c = s[i];
    if (IsDigit(c)) {
      after_comma = false;
      if (blank_seen) return false;
    } else if (c == ',') {
      if (pos == 0 || i + 1 == s.size() || after_comma) return false;
      dot = expo = false;
      after_comma = true;
      blank_seen = false;
    } else {
      after_comma = false;
      if (c == '.') {
        if (dot || expo || blank_seen) return false;
        dot = true;
      } else if (c == 'e' || c == 'E') {
        if (pos == 0 || expo || blank_seen) return false;
        expo = true;
      } else if (c == '-' || c == '+') {
        if ((pos != 0 && s[i - 1] != 'e' && s[i - 1] != 'E') || blank_seen) return false;
      } else {
        return false;
      }
    }
    pos = (c == ',') ? 0 : pos + 1;
    ++i;
    while (i < s.size() && s[i] == ' ') {
      ++i;
      if (c != ',') blank_seen = true;
    }
  }
  return true;
}

// Comma-separated "(re,im)" pairs (parameter_input.cc:622-719).  Numbers
// live only inside parentheses; `digits` counts digits since the last '(',
// ',' or exponent mark and gates signs, exponent marks and commas.
bool ValidComplexList(const std::string& s, bool* ok) {
  size_t i;
  *ok = SkipBlanks(s, &i);
  if (!*ok) return false;
  bool dot = false, expo = false, after_comma = false, blank_seen = false;
  bool open = false;
  int digits = -1;
  while (i < s.size()) {
    const char c = s[i];
    if (IsDigit(c)) {
      ++digits;
      if (blank_seen || !open) return false;
      after_comma = false;
      blank_seen = false;
    } else {
      switch (c) {
        case '.':
          if (dot || expo || blank_seen || !open) return false;
          dot = true;
          break;
        case 'e':
        case 'E':
          if (digits == 0 || expo || blank_seen || !open) return false;
          expo = true;
          digits = 0;
          break;
        case '-':
        case '+':
          if (digits != 0 || blank_seen || !open) return false;
          break;
        case ',':
          if (digits == 0 || i + 1 == s.size() || after_comma) return false;
          dot = expo = false;
          after_comma = true;
          digits = 0;
          break;
        case '(':
          if ((digits > 0 && !after_comma) || open) return false;
          open = true;
          digits = 0;
          break;
        case ')':
          if (!open || digits == 0) return false;
          open = false;
          break;
        default:
          return false;
      }
      if (c != ',') after_comma = false;
      blank_seen = false;
    }
    ++i;
    while (i < s.size() && s[i] == ' ') {
      ++i;
      if (c != ',' && c != '(' && c != ')') blank_seen = true;
    }
  }
  return true;
}

}  // namespace prm

namespace {

// `stream >> value` on one field; a field the stream cannot read gives 0.
double ReadDouble(const std::string& field) {
  std::istringstream in(field);
  double v = 0;
  if (!(in >> v)) v = 0;
  return v;
}

int ReadInt(const std::string& field) {
  std::istringstream in(field);
  int v = 0;
  if (!(in >> v)) v = 0;
  return v;
}

int CountChar(const std::string& s, char c) {
  int n = 0;
  for (char x : s) n += (x == c);
  return n;
}

void Complain(const std::string& what) {
  std::cerr << "ERROR: " << what << "\n";
}

bool GetInt(const std::string& v, const std::string& name, int* out, Range r) {
  bool ok;
  if (!prm::ValidInteger(v, &ok)) {
    std::cerr << "get() Error reading " << name << "\n";
    return false;
  }
  *out = ReadInt(v);
  if (r == kPositive && *out <= 0) { std::cerr << name << " must be positive.\n"; return false; }
  if (r == kNonNegative && *out < 0) { std::cerr << name << " must be non-negative.\n"; return false; }
  return true;
}

bool GetFloat(const std::string& v, const std::string& name, double* out, Range r) {
  bool ok;
  if (!prm::ValidFloat(v, &ok)) {
    std::cerr << "get() Error reading " << name << "\n";
    return false;
  }
  *out = ReadDouble(v);
  if (r == kPositive && *out <= 0) { std::cerr << name << " must be positive.\n"; return false; }
  if (r == kNonNegative && *out < 0) { std::cerr << name << " must be non-negative.\n"; return false; }
  return true;
}

bool GetBool(const std::string& v, const std::string& name, bool* out) {
  bool ok;
  if (!prm::ValidBool(v, &ok)) {
    std::cerr << "get() Error reading " << name << ". Boolean parameters must be 0 or 1.\n";
    return false;
  }
  *out = ReadInt(v) == 1;
  return true;
}

// `size` comma-separated reals into out[0], out[stride], ...
bool GetReals(const std::string& line, int size, int stride, double* out) {
  bool ok;
  if (!prm::ValidRealList(line, &ok)) return false;
  if (!line.empty() && line.back() == ',') return false;
  if (CountChar(line, ',') != size - 1) return false;
  size_t start = 0;
  int n = 0;
  while (true) {
    const size_t comma = line.find(',', start);
    if (comma != std::string::npos && comma <= start) return false;
    const std::string field =
        line.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
    out[(size_t)n * stride] = ReadDouble(field);
    ++n;
    if (comma == std::string::npos) break;
    start = comma + 1;
  }
  return n == size;
}

bool GetInts(const std::string& line, int size, int* out) {
  std::vector<double> tmp(size);
  bool ok;
  if (!prm::ValidRealList(line, &ok)) return false;
  if (!line.empty() && line.back() == ',') return false;
  if (CountChar(line, ',') != size - 1) return false;
  size_t start = 0;
  int n = 0;
  while (true) {
    const size_t comma = line.find(',', start);
    if (comma != std::string::npos && comma <= start) return false;
    out[n++] = ReadInt(line.substr(
        start, comma == std::string::npos ? std::string::npos : comma - start));
    if (comma == std::string::npos) break;
    start = comma + 1;
  }
  return n == size;
}

// `size` complex values, either all real ("1,2") or all "(re,im)", into
// out[0], out[stride], ... (parameter_input.cc:622-719).
bool GetComplexes(const std::string& line, int size, int stride, Complex* out) {
  if (!line.empty() && line.back() == ',') return false;
  const int commas = CountChar(line, ',');
  if (commas != 2 * size - 1 && commas != size - 1) return false;
  const int opens = CountChar(line, '(');
  if (opens != 0 && opens != size) return false;
  if (opens == 0) {
    if (commas != size - 1) return false;
    std::vector<double> re(size);
    if (!GetReals(line, size, 1, re.data())) return false;
    for (int j = 0; j < size; ++j) out[(size_t)j * stride] = Complex(re[j], 0);
    return true;
  }
  bool ok;
  if (!prm::ValidComplexList(line, &ok)) return false;
  int n = 0;
  size_t lp = line.find('(');
  size_t cm = line.find(',', lp);
  size_t rp = line.find(')', cm);
  while (lp != std::string::npos && cm != std::string::npos) {
    if (cm - lp <= 1 && rp - cm <= 1) return false;
    const double re = ReadDouble(line.substr(lp + 1, cm - lp - 1));
    const double im = ReadDouble(line.substr(cm + 1, rp - cm - 1));
    if (n < size) out[(size_t)n * stride] = Complex(re, im);
    ++n;
    lp = line.find('(', rp);
    if (lp == std::string::npos) break;
    cm = line.find(',', lp + 1);
    rp = line.find(')', cm);
  }
  return n == size;
}

// Splits the file the way repeated std::getline + eof() sees it: every
// segment before a '\n', then the trailing segment; eof is raised by the
// read that returns the trailing segment.
class LineSource {
 public:
  explicit LineSource(const std::string& text) {
    size_t start = 0;
    while (true) {
      const size_t nl = text.find('\n', start);
      if (nl == std::string::npos) {
        segs_.push_back(text.substr(start));
        break;
      }
      segs_.push_back(text.substr(start, nl - start));
      start = nl + 1;
    }
  }
  std::string Next() {
    const size_t i = pos_++;
    if (i + 1 >= segs_.size()) eof_ = true;
    return i < segs_.size() ? segs_[i] : std::string();
  }
  bool eof() const { return eof_; }

 private:
  std::vector<std::string> segs_;
  size_t pos_ = 0;
  bool eof_ = false;
};

void FormatComplex(std::ostream& o, const Complex& z) { o << z; }

}  // namespace

PhiParameters::PhiParameters() { SetDefaults(); }

void PhiParameters::SetDefaults() {
  verbose = 1;
  hbar = kHbar;
  boltzmann_constant = kBoltzmann;
  num_states = 0;
  hilbert_space_size = 0;
  hilbert_space_num_elements = 0;
  num_bath_couplings = 0;
  num_matsubara_terms = 1;
  hierarchy_truncation_level = 1;
  matrix_count = 0;
  time_indepenent_hamiltonian.clear();
  initial_annealing_hamiltonian.clear();
  final_annealing_hamiltonian.clear();
  initial_hamiltonian_field.clear();
  final_hamiltonian_field.clear();
  minimum_field_change = 0.01;
  temperature = 300;
  kappa.clear();
  gamma.clear();
  lambda.clear();
  bath_coupling_op.clear();
  multiple_independent_bath_indices.clear();
  initial_density_matrix.clear();
  spectrum_initial_density_matrix.clear();
  integration_time = 0;
  timestep = 0.001;
  output_minimum_timestep = 0.01;
  start_time = 0;
  restart_output_time = 0;
  rkf45_minimum_timestep = 1e-6;
  rkf45_tolerance = 1e-6;
  bicgstab_tolerance = 1e-6;
  filter_tolerance = 0;
  output_filename.clear();
  restart_input_filename.clear();
  restart_output_filename.clear();
  restart_backup_filename.clear();
  is_truncated_system = false;
  use_time_local_truncation = false;
  is_multiple_independent_baths = false;
  is_spectrum_calculation = false;
  is_diagonal_bath_coupling = false;
  is_full_bath_coupling = false;
  is_rho_normalized = true;
  is_restarted = false;
  bicgstab_ell = 2;
  num_threads = 1;
  stupid_hierarchy_partition = false;
  are_parameters_valid = true;
  run_method = NONE;
  bulk_precision = kBulkFp64;
  full_precision_output.clear();
  bicgstab_max_iterations = 20000;
}

void PhiParameters::ReadCommandLine(int num_args, const char* const argv[]) {
  if (num_args < 2) {
    std::cout << "Usage: \n\n  phi paramfile {integrator [threads],print}\n"
              << "  integrator  - One of rk4, rkf45, rk4spectrum, rkf45spectrum, "
              << "steadystate.\n"
              << "  threads     - Number of threads\n"
              << "  memory      - Prints memory requirements\n"
              << "  print       - Prints hierarchy connections\n";
    are_parameters_valid = false;
    return;
  }
  std::cout << "Reading parameters from " << argv[1] << "\n\n";
  {
    std::ifstream probe(argv[1]);
    if (!probe.is_open()) {
      are_parameters_valid = false;
      std::cout << "ERROR: No parameter file.\n";
      return;
    }
  }
  if (num_args == 2) {
    std::cout << "No integration method specified.\n";
    return;
  }
  if (num_args == 4) num_threads = ReadInt(argv[3]);
  const std::string m = argv[2];
  if (m == "rk4spectrum") run_method = RK4SPECTRUM;
  else if (m == "rkf45spectrum") run_method = RKF45SPECTRUM;
  else if (m == "rk4") run_method = RK4;
  else if (m == "rkf45") run_method = RKF45;
  else if (m == "print") run_method = PRINTHIERARCHY;
  else if (m == "memory") run_method = PRINTMEMORY;
  else if (m == "bicgstabl" || m == "bicgstab") run_method = BICGSTABL;
  else if (m == "bicgstabu") run_method = BICGSTABU;
  ReadInputFile(argv[1]);
}

void PhiParameters::ReadInputFile(const std::string& parameter_file) {
  std::string text;
  {
    std::ifstream in(parameter_file.c_str(), std::ios::binary);
    std::ostringstream ss;
    ss << in.rdbuf();
    text = ss.str();
  }
  LineSource src(text);
  std::string line = src.Next(), name, value;
  int row = 0;

  // Allocates a matrix on its first row; returns false on misuse.
  auto matrix_row = [&](std::vector<Complex>* mat, const char* what) -> bool {
    const int hs = hilbert_space_size;
    if (row == 0) {
      mat->assign((size_t)hs * hs, Complex(0, 0));
    } else if (row >= hs) {
      Complain(std::string("Overflow in ") + what + " input.");
      return false;
    }
    if (!GetComplexes(value, hs, hs, mat->data() + row)) {
      Complain(std::string("incorrect ") + name + " input format.");
      return false;
    }
    ++row;
    return true;
  };

  while (!src.eof() && are_parameters_valid) {
    while (line.find('#') != std::string::npos) {
      value.clear();
      line = src.Next();
    }
    const size_t eq = line.find('=');
    const size_t colon = line.find(':');
    if (eq != std::string::npos) {
      name = line.substr(0, eq);
      value = line.substr(eq + 1);
    } else if (colon != std::string::npos) {
      name = line.substr(0, colon);
      row = 0;
      value = src.Next();
    } else if (!line.empty()) {
      value = line;
    } else {
      name.clear();
    }
    line = src.Next();

    bool& ok = are_parameters_valid;
    const std::string& n = name;
    if (n == "fout" || n == "Output" || n == "OutputFile") {
      output_filename = value;
    } else if (n == "Ns" || n == "NumStates") {
      ok = GetInt(value, n, &num_states, kPositive);
      if (!ok) break;
      if (hilbert_space_size == 0) {
        hilbert_space_size = num_states;
        hilbert_space_num_elements = num_states * num_states;
      }
      if (num_bath_couplings == 0) num_bath_couplings = num_states;
      multiple_independent_bath_indices.resize(num_bath_couplings);
      for (int i = 0; i < num_bath_couplings; ++i) multiple_independent_bath_indices[i] = i;
    } else if (n == "HilbertSpaceSize") {
      if (!time_indepenent_hamiltonian.empty() || !final_annealing_hamiltonian.empty()) {
        Complain("HilbertSpaceSize must be specified before the Hamiltonian.");
        ok = false;
        break;
      }
      if (is_diagonal_bath_coupling || is_full_bath_coupling) {
        Complain("HilbertSpaceSize must be specified before the bath coupling operators.");
        ok = false;
        break;
      }
      ok = GetInt(value, n, &hilbert_space_size, kPositive);
      hilbert_space_num_elements = hilbert_space_size * hilbert_space_size;
      if (hilbert_space_size < num_states) {
        Complain("HilbertSpaceSize cannot be less than NumStates.");
        ok = false;
      }
    } else if (n == "M" || n == "NumCouplingTerms") {
      ok = GetInt(value, n, &num_bath_couplings, kPositive);
      if (!ok) break;
      if (num_states != 0) {
        multiple_independent_bath_indices.resize(num_bath_couplings);
        for (int i = 0; i < num_bath_couplings; ++i) multiple_independent_bath_indices[i] = i;
      }
    } else if (n == "Kt" || n == "MatsubaraTerms") {
      ok = GetInt(value, n, &num_matsubara_terms, kNonNegative);
      if (!ok) break;
      num_matsubara_terms += 1;
    } else if (n == "Lh" || n == "HierarchyTruncation") {
      ok = GetInt(value, n, &hierarchy_truncation_level, kPositive);
    } else if (n == "dt" || n == "Timestep") {
      ok = GetFloat(value, n, &timestep, kPositive);
    } else if (n == "t" || n == "Time") {
      ok = GetFloat(value, n, &integration_time, kPositive);
    } else if (n == "OutputMinimumTimestep") {
      ok = GetFloat(value, n, &output_minimum_timestep, kPositive);
    } else if (n == "filter" || n == "filtered" || n == "FilterTolerance") {
      ok = GetFloat(value, n, &filter_tolerance, kPositive);
    } else if (n == "TL" || n == "TimeLocal") {
      ok = GetBool(value, n, &use_time_local_truncation);
    } else if (n == "rho_normalized" || n == "rho_normalised") {
      ok = GetBool(value, n, &is_rho_normalized);
    } else if (n == "H" || n == "Hamiltonian") {
      if (num_states <= 0) {
        Complain("NumStates must be specified before Hamiltonian.");
        ok = false;
        break;
      }
      if (!initial_annealing_hamiltonian.empty() || !final_annealing_hamiltonian.empty() ||
          (!time_indepenent_hamiltonian.empty() && row == 0)) {
        Complain("Hamiltonian already specified");
        ok = false;
        break;
      }
      if (!matrix_row(&time_indepenent_hamiltonian, "Hamiltonian")) { ok = false; break; }
    } else if (n == "H_X" || n == "TransverseHamiltonian") {
      if (num_states <= 0 || row >= hilbert_space_size) { ok = false; break; }
      if (!matrix_row(&initial_annealing_hamiltonian, "TransverseHamiltonian")) { ok = false; break; }
    } else if (n == "TransverseField") {
      const int count = CountChar(value, ',') + 1;
      initial_hamiltonian_field.assign(count, 0);
      if (!GetReals(value, count, 1, initial_hamiltonian_field.data())) {
        Complain("incorrect TransverseField input format.");
        ok = false;
        break;
      }
    } else if (n == "H_Z" || n == "LongitudinalHamiltonian") {
      if (num_states <= 0 || row >= hilbert_space_size) { ok = false; break; }
      if (!matrix_row(&final_annealing_hamiltonian, "LongitudinalHamiltonian")) { ok = false; break; }
    } else if (n == "LongitudinalField") {
      const int count = CountChar(value, ',') + 1;
      final_hamiltonian_field.assign(count, 0);
      if (!GetReals(value, count, 1, final_hamiltonian_field.data())) {
        Complain("incorrect LongitudinalField input format.");
        ok = false;
        break;
      }
    } else if (n == "MinimumFieldChange") {
      ok = GetFloat(value, n, &minimum_field_change, kNonNegative);
    } else if (n == "rho0" || n == "InitialDensityMatrix") {
      if (num_states <= 0) {
        Complain("Ns must be specified before InitialDensityMatrix.");
        ok = false;
        break;
      }
      if (!initial_density_matrix.empty() && row == 0) {
        Complain("InitialDensityMatrix already specified");
        ok = false;
        break;
      }
      if (!matrix_row(&initial_density_matrix, "InitialDensityMatrix")) { ok = false; break; }
    } else if (n == "spectrum_rho0" && spectrum_initial_density_matrix.empty()) {
      if (num_states <= 0) { ok = false; break; }
      spectrum_initial_density_matrix.assign(hilbert_space_size, Complex(0, 0));
      if (!GetComplexes(value, hilbert_space_size, 1, spectrum_initial_density_matrix.data()))
        ok = false;
    } else if (n == "Multibath" || n == "multiBath" || n == "DiagonalCouplingIndices") {
      if (num_bath_couplings <= 0) { ok = false; break; }
      multiple_independent_bath_indices.assign(num_bath_couplings, 0);
      if (!GetInts(value, num_bath_couplings, multiple_independent_bath_indices.data())) {
        ok = false;
        break;
      }
    } else if (n == "Vbath" || n == "DiagonalCouplingTerms" || n == "CorrelatedCouplingTerms") {
      if (num_bath_couplings <= 0 || num_states <= 0) { ok = false; break; }
      if (!bath_coupling_op.empty() && row == 0) { ok = false; break; }
      if (row == 0) {
        is_diagonal_bath_coupling = true;
        bath_coupling_op.assign((size_t)num_bath_couplings * hilbert_space_size, 0);
      } else if (row >= num_bath_couplings) {
        ok = false;
        break;
      }
      if (!GetReals(value, hilbert_space_size, 1,
                    bath_coupling_op.data() + (size_t)hilbert_space_size * row)) {
        ok = false;
        break;
      }
      ++row;
    } else if (n == "FullBathTerms") {
      if (num_bath_couplings <= 0 || num_states <= 0) { ok = false; break; }
      if (!bath_coupling_op.empty() && row == 0) { ok = false; break; }
      if (row == 0) {
        is_full_bath_coupling = true;
        bath_coupling_op.assign((size_t)num_bath_couplings * hilbert_space_num_elements, 0);
      } else if (row >= num_bath_couplings) {
        ok = false;
        break;
      }
      if (!GetReals(value, hilbert_space_num_elements, 1,
                    bath_coupling_op.data() + (size_t)hilbert_space_num_elements * row)) {
        ok = false;
        break;
      }
      ++row;
    } else if (n == "RestartInput") {
      std::istringstream in(value);
      in >> restart_input_filename;
      is_restarted = true;
    } else if (n == "lambda" || n == "Lambda") {
      if (num_bath_couplings <= 0 || !lambda.empty()) { ok = false; break; }
      lambda.assign(num_bath_couplings, 0);
      if (!GetReals(value, num_bath_couplings, 1, lambda.data())) { ok = false; break; }
    } else if (n == "gamma" || n == "Gamma") {
      if (num_bath_couplings <= 0 || !gamma.empty()) { ok = false; break; }
      gamma.assign(num_bath_couplings, 0);
      if (!GetReals(value, num_bath_couplings, 1, gamma.data())) { ok = false; break; }
    } else if (n == "kappa") {
      if (num_states <= 0 || !kappa.empty()) { ok = false; break; }
      kappa.assign(hilbert_space_size, 0);
      if (!GetReals(value, hilbert_space_size, 1, kappa.data())) { ok = false; break; }
    } else if (n == "T" || n == "Temperature") {
      if (!GetFloat(value, n, &temperature, kNonNegative)) { ok = false; break; }
    } else if (n == "verbose") {
      verbose = ReadInt(value);
    } else if (n == "RestartFile") {
      std::istringstream in(value);
      in >> restart_output_filename;
    } else if (n == "RestartOutputTimeSecs") {
      if (!GetFloat(value, n, &restart_output_time, kPositive)) { ok = false; break; }
    } else if (n == "tol_low" || n == "RKF45mindt") {
      if (!GetFloat(value, n, &rkf45_minimum_timestep, kPositive)) { ok = false; break; }
      // PHI applies the steady-state alias only inside its verbose branch
      // (phi_parameters.cc:686-694).
      if (verbose > 0 && n == "tol_low") bicgstab_tolerance = rkf45_minimum_timestep;
    } else if (n == "tol_high" || n == "RKF45tolerance") {
      if (!GetFloat(value, n, &rkf45_tolerance, kPositive)) { ok = false; break; }
    } else if (n == "bicgstab_tolerance") {
      if (!GetFloat(value, n, &bicgstab_tolerance, kPositive)) { ok = false; break; }
    } else if (n == "bicgstab") {
      if (!GetInt(value, n, &bicgstab_ell, kNonNegative)) { ok = false; break; }
    } else if (n == "CyclicPartition") {
      if (!GetBool(value, n, &stupid_hierarchy_partition)) { ok = false; break; }
    } else if (n == "hbar") {
      if (!GetFloat(value, n, &hbar, kPositive)) { ok = false; break; }
    } else if (n == "kB" || n == "k_B") {
      if (!GetFloat(value, n, &boltzmann_constant, kPositive)) { ok = false; break; }
    } else if (n == "BulkPrecision") {
      std::istringstream in(value);
      std::string w;
      in >> w;
      if (w == "fp32") bulk_precision = kBulkFp32;
      else if (w == "fp64") bulk_precision = kBulkFp64;
      else { Complain("BulkPrecision must be fp64 or fp32."); ok = false; break; }
    } else if (n == "FullPrecisionOutput") {
      std::istringstream in(value);
      in >> full_precision_output;
    } else if (n == "BicgstabMaxIterations") {
      if (!GetInt(value, n, &bicgstab_max_iterations, kPositive)) { ok = false; break; }
    }
  }

  if (!are_parameters_valid) {
    std::cout << "Invalid parameters.\n";
    return;
  }
  if (run_method != PRINTMEMORY) FixIncompatibleCombinations();
  matrix_count = GetHierarchyNodeCountEstimate(hierarchy_truncation_level,
                                               num_bath_couplings * num_matsubara_terms);
}

void PhiParameters::FixIncompatibleCombinations() {
  are_parameters_valid =
      !(time_indepenent_hamiltonian.empty() &&
        (initial_annealing_hamiltonian.empty() || final_annealing_hamiltonian.empty()));
  if (!are_parameters_valid && run_method != NONE) {
    std::cout << "Invalid parameters, no Hamiltonian?.\n";
    run_method = BROKENINPUT;
    return;
  }
  if (!initial_annealing_hamiltonian.empty() && initial_hamiltonian_field.empty())
    initial_hamiltonian_field = {1.0, 0.0};
  if (!initial_annealing_hamiltonian.empty() && final_hamiltonian_field.empty())
    final_hamiltonian_field = {0.0, 1.0};
  const bool steady = run_method == BICGSTAB || run_method == BICGSTABL || run_method == BICGSTABU;
  if ((run_method == RKF45 || run_method == RKF45SPECTRUM || steady) && !is_rho_normalized)
    is_rho_normalized = true;
  if (steady) use_time_local_truncation = false;
  if (run_method == BICGSTABL && bicgstab_ell < 1) run_method = BICGSTAB;
  if (filter_tolerance > 0 && (is_diagonal_bath_coupling || is_full_bath_coupling))
    filter_tolerance = 0;
  is_spectrum_calculation = run_method == RK4SPECTRUM || run_method == RKF45SPECTRUM;
  if (is_spectrum_calculation && spectrum_initial_density_matrix.empty()) {
    spectrum_initial_density_matrix.assign(hilbert_space_size, Complex(1, 0));
    if (!initial_density_matrix.empty())
      for (int i = 0; i < hilbert_space_size; ++i)
        spectrum_initial_density_matrix[i] = initial_density_matrix[(size_t)i * hilbert_space_size + i];
  }
  if (multiple_independent_bath_indices.empty()) {
    multiple_independent_bath_indices.resize(num_bath_couplings);
    for (int i = 0; i < num_bath_couplings; ++i) multiple_independent_bath_indices[i] = i;
  }
  if (hilbert_space_size < num_states) {
    hilbert_space_size = num_states;
    hilbert_space_num_elements = num_states * num_states;
  } else if (hilbert_space_size > num_states) {
    is_truncated_system = true;
  }
}

// PHI sums (l + terms - 1)! / (l! (terms - 1)!) over l < max_level in
// double and converts the sum to an integer (phi_parameters.cc:1138-1147).
// Beyond 170! the factorials are infinite, inf / inf is NaN, and PHI's
// conversion of a NaN or out-of-range sum is undefined; here such a sum
// saturates at INT_MAX, and the loops stop once the result is known.
int GetHierarchyNodeCountEstimate(int max_level, int terms) {
  auto product = [](int a, int b) {
    double acc = 1;
    for (int i = a; i <= b && std::isfinite(acc); ++i) acc *= i;
    return acc;
  };
  double count = 0;
  for (int l = 0; l < max_level; ++l) {
    count += product(1, l + terms - 1) / product(1, l) / product(1, terms - 1);
    if (!(count <= (double)INT_MAX)) return INT_MAX;
  }
  return (int)count;
}

void PhiParameters::Write(std::ostream* out, int /*num_matrices*/) const {
  std::ostream& o = *out;
  o << "#PHI " << 1.1 << "\n";
  o << "#Parameters:\n";
  o << "#-----------\n";
  if (hbar != kHbar) o << "#ReducedPlanckConstant=" << hbar << "\n";
  if (boltzmann_constant != kBoltzmann) o << "#BoltzmannConstant=" << boltzmann_constant << "\n";
  o << "#NumStates=" << num_states << "\n";
  if (hilbert_space_size != num_states) o << "#HilbertSpaceSize=" << hilbert_space_size << "\n";
  o << "#MatsubaraTerms=" << num_matsubara_terms - 1 << "\n";
  o << "#CouplingTerms=" << num_bath_couplings << "\n";
  o << "#Temperature=" << temperature << "\n";
  o << "#HierarchyTruncation=" << hierarchy_truncation_level << "\n";
  o << "#TimeLocal=" << (use_time_local_truncation ? 1 : 0) << "\n";
  o << "#Timestep=" << timestep << "\n";
  o << "#Time=" << integration_time << "\n";
  o << "#filter=" << filter_tolerance << "\n";
  o << "#RKF45tolerance = " << rkf45_tolerance << "\n";
  o << "#RKF45mindt = " << rkf45_minimum_timestep << "\n";
  auto list = [&](const char* title, const std::vector<Float>& v, size_t count) {
    o << "#" << title << ":\n#";
    for (size_t i = 0; i < count; ++i) o << (i ? "," : "") << v[i];
    o << "\n";
  };
  if (!gamma.empty()) list("gamma", gamma, num_bath_couplings);
  if (!lambda.empty()) list("lambda", lambda, num_bath_couplings);
  if (!kappa.empty()) list("kappa", kappa, num_states);
  if (!initial_hamiltonian_field.empty())
    list("TransverseField", initial_hamiltonian_field, initial_hamiltonian_field.size());
  if (!final_hamiltonian_field.empty())
    list("LongitudinalField", final_hamiltonian_field, final_hamiltonian_field.size());
  if (is_multiple_independent_baths) {
    o << "#multiBath:\n#";
    for (int i = 0; i < num_bath_couplings; ++i)
      o << (i ? "," : "") << multiple_independent_bath_indices[i];
    o << "\n";
  }
  const int hs = hilbert_space_size;
  if (is_diagonal_bath_coupling) {
    o << "#CorrelatedCouplingTerms:\n";
    for (int m = 0; m < num_bath_couplings; ++m) {
      o << "#";
      for (int j = 0; j < hs; ++j) o << (j ? "," : "") << bath_coupling_op[(size_t)m * hs + j];
      o << "\n";
    }
  }
  if (is_full_bath_coupling) {
    const int ne = hilbert_space_num_elements;
    o << "#FullBathTerms:\n";
    for (int m = 0; m < num_bath_couplings; ++m) {
      o << "#";
      for (int j = 0; j < ne; ++j) o << (j ? "," : "") << bath_coupling_op[(size_t)m * ne + j];
      o << "\n";
    }
  }
  // A matrix prints row i as elements (i, 0..hs-2) followed by column-major
  // element (i+1)*hs - 1, which is (hs-1, i): PHI's writer
  // (phi_parameters.cc:1047-1090) prints the transpose's entry last.
  auto matrix = [&](const char* title, const std::vector<Complex>& a) {
    o << "#" << title << ":\n";
    for (int i = 0; i < hs; ++i) {
      o << "#";
      for (int j = 0; j < hs - 1; ++j) {
        FormatComplex(o, a[(size_t)i + (size_t)j * hs]);
        o << ",";
      }
      FormatComplex(o, a[(size_t)(i + 1) * hs - 1]);
      o << "\n";
    }
  };
  if (!initial_annealing_hamiltonian.empty()) matrix("TransverseHamiltonian", initial_annealing_hamiltonian);
  if (!final_annealing_hamiltonian.empty()) matrix("LongitudinalHamiltonian", final_annealing_hamiltonian);
  if (!time_indepenent_hamiltonian.empty()) matrix("Hamiltonian", time_indepenent_hamiltonian);
  if (!initial_density_matrix.empty()) matrix("InitialDensityMatrix", initial_density_matrix);
  if (!restart_input_filename.empty()) o << "#RestartInput=" << restart_input_filename << "\n";
  if (restart_output_time > 0) {
    o << "#RestartOutputTime=" << restart_output_time << "\n";
    o << "#RestartFile=" << restart_output_filename << "\n";
  }
  o << "#Estimated Number of Matrices="
    << GetHierarchyNodeCountEstimate(hierarchy_truncation_level,
                                     num_bath_couplings * num_matsubara_terms)
    << "\n";
  o << "#-------------\n";
  const char* integrator = nullptr;
  switch (run_method) {
    case RK4: integrator = "rk4"; break;
    case RKF45: integrator = "rkf45"; break;
    case RK4SPECTRUM: integrator = "rk4spectrum"; break;
    case RKF45SPECTRUM: integrator = "rkf45spectrum"; break;
    default: break;
  }
  if (integrator) {
    o << "#Threads=" << num_threads << "\n";
    o << "#Integrator=" << integrator << "\n";
  } else if (run_method == BICGSTAB || run_method == BICGSTABU || run_method == BICGSTABL) {
    o << "#Threads=" << num_threads << "\n";
    if (run_method == BICGSTABL)
      o << "#Calculating steady-state using BiCGSTAB(" << bicgstab_ell << ")\n";
    else
      o << "#Calculating steady-state using BiCGSTAB\n";
  }
}

}  // namespace phi
